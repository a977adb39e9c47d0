// Host-side graph construction: seeded R-MAT / Erdos-Renyi generators and the
// CSR views the device path uploads (out, in, all = in++out, symmetrised).
//
// The reference builds its adjacency by one sorted insert per edge
// (CPU_adj_list.hpp:110-122, O(deg) each) and flattens it in toCSR
// (GPU_csr.hpp:282-339).  Here the same final state -- per-vertex neighbour
// lists ascending, duplicates merged with the last weight winning -- is built
// by counting sort + per-row sort, threaded, so scale-24 graphs build in
// seconds.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <atomic>
#include <cerrno>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>

#include "host_common.hpp"

namespace ghet {

static thread_local std::string g_err;

int set_error(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}
void clear_error() { g_err.clear(); }
const char* last_error_cstr() { return g_err.c_str(); }

int64_t exclusive_scan_i64(std::vector<int64_t>& v, int threads) {
    const int64_t n = static_cast<int64_t>(v.size());
    threads = std::max(1, std::min<int>(threads, static_cast<int>(n / 65536 + 1)));
    std::vector<int64_t> part(threads + 1, 0);
    parallel_for(n, threads, [&](int64_t b, int64_t e, int t) {
        int64_t s = 0;
        for (int64_t i = b; i < e; ++i) s += v[i];
        part[t + 1] = s;
    });
    for (int t = 0; t < threads; ++t) part[t + 1] += part[t];
    parallel_for(n, threads, [&](int64_t b, int64_t e, int t) {
        int64_t s = part[t];
        for (int64_t i = b; i < e; ++i) {
            int64_t x = v[i];
            v[i] = s;
            s += x;
        }
    });
    return part[threads];
}

}  // namespace ghet

using namespace ghet;

struct ghet_host_graph {
    int32_t n = 0;
    int threads = 1;
    std::vector<int32_t> out_rowptr, out_col;
    std::vector<double> out_w;
    bool have_in = false, have_all = false, have_sym = false;
    std::vector<int32_t> in_rowptr, in_col;
    std::vector<double> in_w;
    std::vector<int32_t> all_rowptr, all_col;
    std::vector<int32_t> sym_rowptr, sym_col;
    std::mutex mu;
};

namespace {

// Builds the out-CSR from `keys` bucketed by source.  keys[i] = (v << 32) | k,
// where k is the sample index; per row, entries are sorted and the largest k
// of each v is kept (the last sample wins).  weight_of(k) gives the weight.
template <typename WeightOf>
int finish_rows(ghet_host_graph* g, std::vector<int64_t>& offs, std::vector<uint64_t>& keys,
                WeightOf&& weight_of) {
    const int32_t n = g->n;
    const int T = g->threads;
    std::vector<int64_t> deg(static_cast<size_t>(n) + 1, 0);
    parallel_for(n, T, [&](int64_t b, int64_t e, int) {
        for (int64_t u = b; u < e; ++u) {
            uint64_t* row = keys.data() + offs[u];
            int64_t len = offs[u + 1] - offs[u];
            std::sort(row, row + len);
            int64_t out = 0;
            for (int64_t i = 0; i < len; ++i) {
                // keep the last entry of each run of equal v (largest k)
                if (i + 1 < len && (row[i + 1] >> 32) == (row[i] >> 32)) continue;
                row[out++] = row[i];
            }
            deg[u] = out;
        }
    });
    int64_t m = exclusive_scan_i64(deg, T);
    if (m > INT32_MAX) return set_error(GHET_ERR_UNSUPPORTED, "graph has %lld edges (> 2^31-1)", (long long)m);
    g->out_rowptr.resize(static_cast<size_t>(n) + 1);
    g->out_col.resize(static_cast<size_t>(m));
    g->out_w.resize(static_cast<size_t>(m));
    parallel_for(n, T, [&](int64_t b, int64_t e, int) {
        for (int64_t u = b; u < e; ++u) {
            const uint64_t* row = keys.data() + offs[u];
            int64_t len = (u + 1 <= n ? deg[u + 1] : m) - deg[u];
            int64_t dst = deg[u];
            g->out_rowptr[u] = static_cast<int32_t>(dst);
            for (int64_t i = 0; i < len; ++i) {
                g->out_col[dst + i] = static_cast<int32_t>(row[i] >> 32);
                g->out_w[dst + i] = weight_of(static_cast<uint32_t>(row[i] & 0xFFFFFFFFu));
            }
        }
    });
    g->out_rowptr[n] = static_cast<int32_t>(m);
    return GHET_OK;
}

// Bucket (u, key) pairs by u: pairs[i] = (u << 32) | v for sample i (or ~0 for
// a dropped sample).  Produces offs (n+1) and keys = (v << 32) | i.
int bucket_by_source(ghet_host_graph* g, const std::vector<uint64_t>& pairs,
                     std::vector<int64_t>& offs, std::vector<uint64_t>& keys) {
    const int32_t n = g->n;
    const int T = g->threads;
    const int64_t S = static_cast<int64_t>(pairs.size());
    std::unique_ptr<std::atomic<int64_t>[]> cnt(new (std::nothrow) std::atomic<int64_t>[static_cast<size_t>(n) + 1]);
    if (!cnt) return set_error(GHET_ERR_NOMEM, "host allocation failed (%d counters)", n);
    parallel_for(static_cast<int64_t>(n) + 1, T, [&](int64_t b, int64_t e, int) {
        for (int64_t i = b; i < e; ++i) cnt[i].store(0, std::memory_order_relaxed);
    });
    parallel_for(S, T, [&](int64_t b, int64_t e, int) {
        for (int64_t i = b; i < e; ++i) {
            if (pairs[i] == ~0ULL) continue;
            cnt[pairs[i] >> 32].fetch_add(1, std::memory_order_relaxed);
        }
    });
    offs.assign(static_cast<size_t>(n) + 1, 0);
    for (int64_t i = 0; i < n; ++i) offs[i] = cnt[i].load(std::memory_order_relaxed);
    int64_t total = exclusive_scan_i64(offs, T);
    offs[n] = total;
    parallel_for(n, T, [&](int64_t b, int64_t e, int) {
        for (int64_t i = b; i < e; ++i) cnt[i].store(offs[i], std::memory_order_relaxed);
    });
    keys.resize(static_cast<size_t>(total));
    parallel_for(S, T, [&](int64_t b, int64_t e, int) {
        for (int64_t i = b; i < e; ++i) {
            uint64_t p = pairs[i];
            if (p == ~0ULL) continue;
            int64_t pos = cnt[p >> 32].fetch_add(1, std::memory_order_relaxed);
            keys[pos] = ((p & 0xFFFFFFFFULL) << 32) | static_cast<uint64_t>(i);
        }
    });
    return GHET_OK;
}

void build_in(ghet_host_graph* g) {
    const int32_t n = g->n;
    const int T = g->threads;
    const int64_t m = g->out_rowptr[n];
    std::unique_ptr<std::atomic<int32_t>[]> cnt(new std::atomic<int32_t>[static_cast<size_t>(n) + 1]);
    for (int64_t i = 0; i <= n; ++i) cnt[i].store(0, std::memory_order_relaxed);
    parallel_for(m, T, [&](int64_t b, int64_t e, int) {
        for (int64_t i = b; i < e; ++i) cnt[g->out_col[i]].fetch_add(1, std::memory_order_relaxed);
    });
    std::vector<int64_t> offs(static_cast<size_t>(n) + 1);
    for (int64_t i = 0; i < n; ++i) offs[i] = cnt[i].load(std::memory_order_relaxed);
    offs[n] = 0;
    exclusive_scan_i64(offs, T);
    offs[n] = m;
    for (int64_t i = 0; i < n; ++i) cnt[i].store(static_cast<int32_t>(offs[i]), std::memory_order_relaxed);
    std::vector<uint64_t> keys(static_cast<size_t>(m));  // (u << 32) | edge index
    parallel_for(n, T, [&](int64_t b, int64_t e, int) {
        for (int64_t u = b; u < e; ++u)
            for (int32_t i = g->out_rowptr[u]; i < g->out_rowptr[u + 1]; ++i) {
                int32_t pos = cnt[g->out_col[i]].fetch_add(1, std::memory_order_relaxed);
                keys[pos] = (static_cast<uint64_t>(u) << 32) | static_cast<uint32_t>(i);
            }
    });
    g->in_rowptr.resize(static_cast<size_t>(n) + 1);
    g->in_col.resize(static_cast<size_t>(m));
    g->in_w.resize(static_cast<size_t>(m));
    parallel_for(n, T, [&](int64_t b, int64_t e, int) {
        for (int64_t v = b; v < e; ++v) {
            std::sort(keys.begin() + offs[v], keys.begin() + offs[v + 1]);
            g->in_rowptr[v] = static_cast<int32_t>(offs[v]);
            for (int64_t i = offs[v]; i < offs[v + 1]; ++i) {
                g->in_col[i] = static_cast<int32_t>(keys[i] >> 32);
                g->in_w[i] = g->out_w.empty() ? 1.0 : g->out_w[keys[i] & 0xFFFFFFFFULL];
            }
        }
    });
    g->in_rowptr[n] = static_cast<int32_t>(m);
    g->have_in = true;
}

int build_all(ghet_host_graph* g) {
    if (!g->have_in) build_in(g);
    const int32_t n = g->n;
    const int64_t tot = static_cast<int64_t>(g->out_rowptr[n]) + g->in_rowptr[n];
    if (tot > INT32_MAX) return set_error(GHET_ERR_UNSUPPORTED, "all-CSR has %lld entries (> 2^31-1)", (long long)tot);
    g->all_rowptr.resize(static_cast<size_t>(n) + 1);
    g->all_col.resize(static_cast<size_t>(tot));
    parallel_for(n + 1, g->threads, [&](int64_t b, int64_t e, int) {
        for (int64_t v = b; v < e; ++v) g->all_rowptr[v] = g->out_rowptr[v] + g->in_rowptr[v];
    });
    parallel_for(n, g->threads, [&](int64_t b, int64_t e, int) {
        for (int64_t v = b; v < e; ++v) {
            int32_t d = g->all_rowptr[v];
            for (int32_t i = g->in_rowptr[v]; i < g->in_rowptr[v + 1]; ++i) g->all_col[d++] = g->in_col[i];
            for (int32_t i = g->out_rowptr[v]; i < g->out_rowptr[v + 1]; ++i) g->all_col[d++] = g->out_col[i];
        }
    });
    g->have_all = true;
    return GHET_OK;
}

void build_sym(ghet_host_graph* g) {
    if (!g->have_in) build_in(g);
    const int32_t n = g->n;
    std::vector<int64_t> deg(static_cast<size_t>(n) + 1, 0);
    auto merge = [&](int64_t v, int32_t* dst) -> int64_t {
        int32_t i = g->out_rowptr[v], ie = g->out_rowptr[v + 1];
        int32_t j = g->in_rowptr[v], je = g->in_rowptr[v + 1];
        int64_t c = 0;
        int32_t last = -1;
        while (i < ie || j < je) {
            int32_t x;
            if (j >= je || (i < ie && g->out_col[i] <= g->in_col[j])) x = g->out_col[i++];
            else x = g->in_col[j++];
            if (x == v || x == last) continue;
            last = x;
            if (dst) dst[c] = x;
            ++c;
        }
        return c;
    };
    parallel_for(n, g->threads, [&](int64_t b, int64_t e, int) {
        for (int64_t v = b; v < e; ++v) deg[v] = merge(v, nullptr);
    });
    int64_t tot = exclusive_scan_i64(deg, g->threads);
    deg[n] = tot;
    g->sym_rowptr.resize(static_cast<size_t>(n) + 1);
    g->sym_col.resize(static_cast<size_t>(tot));
    parallel_for(n, g->threads, [&](int64_t b, int64_t e, int) {
        for (int64_t v = b; v < e; ++v) {
            g->sym_rowptr[v] = static_cast<int32_t>(deg[v]);
            merge(v, g->sym_col.data() + deg[v]);
        }
    });
    g->sym_rowptr[n] = static_cast<int32_t>(tot);
    g->have_sym = true;
}

}  // namespace

extern "C" {

const char* ghet_last_error(void) { return last_error_cstr(); }
int ghet_abi_version(void) { return GHET_ABI_VERSION; }

int ghet_gen_rmat(int scale, double edge_factor, double a, double b, double c, uint64_t seed, int threads,
                  ghet_host_graph** out) {
    clear_error();
    if (!out) return set_error(GHET_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (scale < 1 || scale > 30) return set_error(GHET_ERR_INVALID, "scale %d outside [1,30]", scale);
    if (!(edge_factor > 0) || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        return set_error(GHET_ERR_INVALID, "bad R-MAT parameters");
    const double samples_d = edge_factor * std::ldexp(1.0, scale);
    if (samples_d >= 4294967295.0) return set_error(GHET_ERR_UNSUPPORTED, "too many samples (%g)", samples_d);
    const int64_t S = static_cast<int64_t>(samples_d);
    auto* g = new (std::nothrow) ghet_host_graph();
    if (!g) return set_error(GHET_ERR_NOMEM, "host allocation failed");
    g->n = static_cast<int32_t>(1LL << scale);
    g->threads = hw_threads(threads);
    // Quadrant thresholds on 32-bit uniforms.
    const uint64_t t1 = static_cast<uint64_t>(a * 4294967296.0);
    const uint64_t t2 = static_cast<uint64_t>((a + b) * 4294967296.0);
    const uint64_t t3 = static_cast<uint64_t>((a + b + c) * 4294967296.0);
    std::vector<uint64_t> pairs;
    try {
        pairs.resize(static_cast<size_t>(S));
    } catch (...) {
        delete g;
        return set_error(GHET_ERR_NOMEM, "host allocation failed (%lld samples)", (long long)S);
    }
    parallel_for(S, g->threads, [&](int64_t bgn, int64_t end, int) {
        for (int64_t k = bgn; k < end; ++k) {
            uint64_t key = stream_key(seed, static_cast<uint64_t>(k));
            uint32_t u = 0, v = 0;
            uint64_t r = 0;
            for (int j = 0; j < scale; ++j) {
                if ((j & 1) == 0) r = draw(key, static_cast<uint64_t>(j >> 1));
                uint64_t x = (j & 1) ? (r >> 32) : (r & 0xFFFFFFFFULL);
                uint32_t q = x < t1 ? 0u : x < t2 ? 1u : x < t3 ? 2u : 3u;
                u = (u << 1) | (q >> 1);
                v = (v << 1) | (q & 1u);
            }
            pairs[k] = (u == v) ? ~0ULL : ((static_cast<uint64_t>(u) << 32) | v);
        }
    });
    std::vector<int64_t> offs;
    std::vector<uint64_t> keys;
    int st = bucket_by_source(g, pairs, offs, keys);
    std::vector<uint64_t>().swap(pairs);
    if (st != GHET_OK) { delete g; return st; }
    const uint64_t wdraw = static_cast<uint64_t>((scale + 1) / 2);
    st = finish_rows(g, offs, keys, [&](uint32_t k) { return weight_from(draw(stream_key(seed, k), wdraw)); });
    if (st != GHET_OK) { delete g; return st; }
    *out = g;
    return GHET_OK;
}

int ghet_gen_er(int32_t n, int64_t m, uint64_t seed, int threads, ghet_host_graph** out) {
    clear_error();
    if (!out) return set_error(GHET_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (n < 2 || m < 0 || m > static_cast<int64_t>(n) * (n - 1) / 2 || m > INT32_MAX)
        return set_error(GHET_ERR_INVALID, "bad ER parameters n=%d m=%lld", n, (long long)m);
    auto* g = new (std::nothrow) ghet_host_graph();
    if (!g) return set_error(GHET_ERR_NOMEM, "host allocation failed");
    g->n = n;
    g->threads = hw_threads(threads);
    // Draw samples in index order until m distinct non-loop pairs exist; the
    // edge set is the distinct pairs among the shortest such prefix.  Each
    // kept pair takes the weight of its first sample.
    int64_t S = m + m / 8 + 1024;
    for (;;) {
        if (S >= 4294967295LL) { delete g; return set_error(GHET_ERR_UNSUPPORTED, "ER sample count overflow"); }
        std::vector<uint64_t> pairs(static_cast<size_t>(S));
        parallel_for(S, g->threads, [&](int64_t bgn, int64_t end, int) {
            for (int64_t k = bgn; k < end; ++k) {
                uint64_t key = stream_key(seed, static_cast<uint64_t>(k));
                uint64_t u = static_cast<uint64_t>((static_cast<unsigned __int128>(draw(key, 0)) * n) >> 64);
                uint64_t v = static_cast<uint64_t>((static_cast<unsigned __int128>(draw(key, 1)) * n) >> 64);
                pairs[k] = (u == v) ? ~0ULL : ((u << 32) | v);
            }
        });
        std::vector<int64_t> offs;
        std::vector<uint64_t> keys;
        int st = bucket_by_source(g, pairs, offs, keys);
        if (st != GHET_OK) { delete g; return st; }
        // first occurrence index of each distinct pair
        std::vector<uint32_t> firsts;
        std::vector<int64_t> cnt(g->threads + 1, 0);
        parallel_for(n, g->threads, [&](int64_t b, int64_t e, int t) {
            int64_t c = 0;
            for (int64_t u = b; u < e; ++u) {
                uint64_t* row = keys.data() + offs[u];
                int64_t len = offs[u + 1] - offs[u];
                std::sort(row, row + len);
                for (int64_t i = 0; i < len; ++i)
                    if (i == 0 || (row[i] >> 32) != (row[i - 1] >> 32)) ++c;
            }
            cnt[t + 1] = c;
        });
        int64_t distinct = 0;
        for (int t = 0; t < g->threads; ++t) distinct += cnt[t + 1];
        if (distinct < m) {
            S = S + (m - distinct) * 2 + 1024;
            continue;
        }
        firsts.reserve(static_cast<size_t>(distinct));
        for (int64_t u = 0; u < n; ++u) {
            const uint64_t* row = keys.data() + offs[u];
            int64_t len = offs[u + 1] - offs[u];
            for (int64_t i = 0; i < len; ++i)
                if (i == 0 || (row[i] >> 32) != (row[i - 1] >> 32)) firsts.push_back(static_cast<uint32_t>(row[i]));
        }
        uint32_t cutoff = 0xFFFFFFFFu;  // keep pairs whose first sample index < cutoff
        if (distinct > m) {
            std::nth_element(firsts.begin(), firsts.begin() + (m - 1), firsts.end());
            cutoff = firsts[m - 1] + 1;
        }
        // keep only the first occurrence of each pair, below the cutoff
        parallel_for(n, g->threads, [&](int64_t b, int64_t e, int) {
            for (int64_t u = b; u < e; ++u) {
                uint64_t* row = keys.data() + offs[u];
                int64_t len = offs[u + 1] - offs[u];
                for (int64_t i = 0; i < len; ++i) {
                    bool first = (i == 0 || (row[i] >> 32) != (row[i - 1] >> 32));
                    if (!first || static_cast<uint32_t>(row[i]) >= cutoff) row[i] |= 0xFFFFFFFFULL;
                }
            }
        });
        // compact: drop marked entries (marker = low 32 bits all ones)
        std::vector<int64_t> offs2(static_cast<size_t>(n) + 1, 0);
        std::vector<uint64_t> keys2;
        keys2.reserve(static_cast<size_t>(m));
        for (int64_t u = 0; u < n; ++u) {
            offs2[u] = static_cast<int64_t>(keys2.size());
            for (int64_t i = offs[u]; i < offs[u + 1]; ++i)
                if ((keys[i] & 0xFFFFFFFFULL) != 0xFFFFFFFFULL) keys2.push_back(keys[i]);
        }
        offs2[n] = static_cast<int64_t>(keys2.size());
        st = finish_rows(g, offs2, keys2, [&](uint32_t k) { return weight_from(draw(stream_key(seed, k), 2)); });
        if (st != GHET_OK) { delete g; return st; }
        break;
    }
    *out = g;
    return GHET_OK;
}

int ghet_host_graph_from_edges(int32_t n, int64_t m, const int32_t* src, const int32_t* dst, const double* w,
                               int threads, ghet_host_graph** out) {
    clear_error();
    if (!out) return set_error(GHET_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (n < 0 || m < 0 || (m > 0 && (!src || !dst))) return set_error(GHET_ERR_INVALID, "bad edge list");
    if (m >= 4294967295LL) return set_error(GHET_ERR_UNSUPPORTED, "too many edges");
    for (int64_t i = 0; i < m; ++i)
        if (src[i] < 0 || src[i] >= n || dst[i] < 0 || dst[i] >= n)
            return set_error(GHET_ERR_INVALID, "edge %lld (%d,%d) out of range [0,%d)", (long long)i, src[i], dst[i], n);
    auto* g = new (std::nothrow) ghet_host_graph();
    if (!g) return set_error(GHET_ERR_NOMEM, "host allocation failed");
    g->n = n;
    g->threads = hw_threads(threads);
    std::vector<uint64_t> pairs(static_cast<size_t>(m));
    for (int64_t i = 0; i < m; ++i) pairs[i] = (static_cast<uint64_t>(src[i]) << 32) | static_cast<uint32_t>(dst[i]);
    std::vector<int64_t> offs;
    std::vector<uint64_t> keys;
    int st = bucket_by_source(g, pairs, offs, keys);
    if (st != GHET_OK) { delete g; return st; }
    if (n == 0) {
        g->out_rowptr.assign(1, 0);
    } else {
        st = finish_rows(g, offs, keys, [&](uint32_t k) { return w ? w[k] : 1.0; });
        if (st != GHET_OK) { delete g; return st; }
    }
    *out = g;
    return GHET_OK;
}

int ghet_host_graph_view(ghet_host_graph* g, int kind, ghet_csr_view* out) {
    clear_error();
    if (!g || !out) return set_error(GHET_ERR_INVALID, "NULL argument");
    std::lock_guard<std::mutex> lock(g->mu);
    switch (kind) {
        case GHET_VIEW_OUT:
            *out = {g->n, g->out_rowptr.empty() ? 0 : g->out_rowptr[g->n], g->out_rowptr.data(), g->out_col.data(),
                    g->out_w.data()};
            return GHET_OK;
        case GHET_VIEW_IN:
            if (!g->have_in) build_in(g);
            *out = {g->n, g->in_rowptr[g->n], g->in_rowptr.data(), g->in_col.data(), g->in_w.data()};
            return GHET_OK;
        case GHET_VIEW_ALL: {
            if (!g->have_all) {
                int st = build_all(g);
                if (st != GHET_OK) return st;
            }
            *out = {g->n, g->all_rowptr[g->n], g->all_rowptr.data(), g->all_col.data(), nullptr};
            return GHET_OK;
        }
        case GHET_VIEW_SYM:
            if (!g->have_sym) build_sym(g);
            *out = {g->n, g->sym_rowptr[g->n], g->sym_rowptr.data(), g->sym_col.data(), nullptr};
            return GHET_OK;
        default:
            return set_error(GHET_ERR_INVALID, "unknown view kind %d", kind);
    }
}

void ghet_host_graph_free(ghet_host_graph* g) { delete g; }

}  // extern "C"

namespace ghet {

ghet_host_graph* host_graph_adopt(int32_t n, const int32_t* rowptr, const int32_t* col, const double* w,
                                  int threads) {
    auto* g = new (std::nothrow) ghet_host_graph();
    if (!g) return nullptr;
    g->n = n;
    g->threads = hw_threads(threads);
    const int64_t m = rowptr[n];
    g->out_rowptr.assign(rowptr, rowptr + n + 1);
    g->out_col.assign(col, col + m);
    if (w) g->out_w.assign(w, w + m);
    else g->out_w.assign(static_cast<size_t>(m), 1.0);
    return g;
}

}  // namespace ghet

// ------------------------------------------------------------ CSR cache ----
// Two files in the vector-of-vectors layout (int32 row count, then per row an
// int32 length and its elements): targets in `path`, weights in `path`.w.
// Row u's elements start at byte 4 + 4(u+1) + sizeof(T) * rowptr[u], so once
// the target file's headers give rowptr, every other offset is known and the
// copies and checks run in parallel.
namespace ghet {
namespace {

struct MappedFile {
    const unsigned char* p = nullptr;
    size_t size = 0;
    int fd = -1;
    ~MappedFile() {
        if (p && size) munmap(const_cast<unsigned char*>(p), size);
        if (fd >= 0) close(fd);
    }
    // 0 ok, 1 missing, -1 other failure
    int open_ro(const std::string& path) {
        fd = ::open(path.c_str(), O_RDONLY);
        if (fd < 0) return errno == ENOENT ? 1 : -1;
        struct stat sb;
        if (fstat(fd, &sb) != 0) return -1;
        size = static_cast<size_t>(sb.st_size);
        if (size == 0) return 0;
        void* m = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
        if (m == MAP_FAILED) {
            size = 0;
            return -1;
        }
        p = static_cast<const unsigned char*>(m);
        return 0;
    }
    int32_t i32(size_t off) const {
        int32_t x;
        std::memcpy(&x, p + off, 4);
        return x;
    }
};

// Rows with unsorted or repeated targets: rebuilt as an edge list in row
// order (sorted, the later entry's weight kept), like
// ghet_host_graph_from_edges.
int rebuild_rows(ghet_host_graph* g) {
    const int32_t n = g->n;
    const int64_t m = g->out_rowptr[n];
    const int32_t* rp = g->out_rowptr.data();
    std::vector<uint64_t> pairs(static_cast<size_t>(m));
    parallel_for(n, g->threads, [&](int64_t b, int64_t e, int) {
        for (int64_t u = b; u < e; ++u)
            for (int32_t k = rp[u]; k < rp[u + 1]; ++k)
                pairs[k] = (static_cast<uint64_t>(u) << 32) | static_cast<uint32_t>(g->out_col[k]);
    });
    std::vector<double> w = std::move(g->out_w);
    std::vector<int64_t> offs;
    std::vector<uint64_t> keys;
    int st = bucket_by_source(g, pairs, offs, keys);
    if (st != GHET_OK) return st;
    pairs.clear();
    pairs.shrink_to_fit();
    return finish_rows(g, offs, keys, [&](uint32_t k) { return w[k]; });
}

template <typename T>
int save_rows(const std::string& path, int32_t n, const int32_t* rowptr, const T* data) {
    const std::string tmp = path + ".tmp." + std::to_string(static_cast<long long>(getpid()));
    std::FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) return set_error(GHET_ERR_IO, "Unable to open file %s", tmp.c_str());
    std::vector<char> buf(size_t(1) << 22);
    std::setvbuf(f, buf.data(), _IOFBF, buf.size());
    bool ok = std::fwrite(&n, 4, 1, f) == 1;
    for (int32_t u = 0; ok && u < n; ++u) {
        const int32_t len = rowptr[u + 1] - rowptr[u];
        ok = std::fwrite(&len, 4, 1, f) == 1;
        if (ok && len > 0) ok = std::fwrite(data + rowptr[u], sizeof(T), static_cast<size_t>(len), f) == size_t(len);
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok || std::rename(tmp.c_str(), path.c_str()) != 0) {
        std::remove(tmp.c_str());
        return set_error(GHET_ERR_IO, "Unable to write file %s", path.c_str());
    }
    return GHET_OK;
}

}  // namespace
}  // namespace ghet

extern "C" {

int ghet_host_graph_save(ghet_host_graph* g, const char* path) {
    clear_error();
    if (!g || !path) return set_error(GHET_ERR_INVALID, "NULL argument");
    const std::string p(path);
    const int32_t* rp = g->out_rowptr.data();
    int st = save_rows<int32_t>(p, g->n, rp, g->out_col.data());
    if (st != GHET_OK) return st;
    return save_rows<double>(p + ".w", g->n, rp, g->out_w.data());
}

int ghet_host_graph_load(const char* path, int threads, ghet_host_graph** out) {
    clear_error();
    if (!path || !out) return set_error(GHET_ERR_INVALID, "NULL argument");
    *out = nullptr;
    const std::string p(path);
    MappedFile tf, wf;
    if (tf.open_ro(p) != 0) return set_error(GHET_ERR_IO, "Unable to open file %s", path);
    const int wst = wf.open_ro(p + ".w");
    if (wst < 0) return set_error(GHET_ERR_IO, "Unable to open file %s.w", path);
    const bool have_w = wst == 0;
    if (tf.size < 4) return set_error(GHET_ERR_IO, "Unable to read file %s (no row count)", path);
    const int32_t n = tf.i32(0);
    if (n < 0) return set_error(GHET_ERR_IO, "Unable to read file %s (row count %d)", path, n);
    // headers of the target file: row lengths -> rowptr
    std::vector<int32_t> rowptr(static_cast<size_t>(n) + 1, 0);
    size_t pos = 4;
    int64_t m = 0;
    for (int32_t u = 0; u < n; ++u) {
        if (pos + 4 > tf.size) return set_error(GHET_ERR_IO, "Unable to read file %s (row %d header)", path, u);
        const int32_t len = tf.i32(pos);
        if (len < 0 || pos + 4 + 4 * size_t(len) > tf.size)
            return set_error(GHET_ERR_IO, "Unable to read file %s (row %d length %d)", path, u, len);
        m += len;
        if (m > INT32_MAX) return set_error(GHET_ERR_UNSUPPORTED, "graph has more than 2^31-1 edges");
        rowptr[u + 1] = static_cast<int32_t>(m);
        pos += 4 + 4 * size_t(len);
    }
    if (pos != tf.size) return set_error(GHET_ERR_IO, "Unable to read file %s (%zu trailing bytes)", path, tf.size - pos);
    if (have_w) {
        const size_t want = 4 + 4 * size_t(n) + 8 * size_t(m);
        if (wf.size != want || (wf.size >= 4 && wf.i32(0) != n))
            return set_error(GHET_ERR_IO, "Unable to read file %s.w (does not match %s)", path, path);
    }
    auto* g = new (std::nothrow) ghet_host_graph();
    if (!g) return set_error(GHET_ERR_NOMEM, "host allocation failed");
    std::unique_ptr<ghet_host_graph> guard(g);
    g->n = n;
    g->threads = hw_threads(threads);
    const int T = g->threads;
    g->out_rowptr = std::move(rowptr);
    const int32_t* rp = g->out_rowptr.data();
    g->out_col.resize(static_cast<size_t>(m));
    g->out_w.resize(static_cast<size_t>(m));
    std::atomic<int> bad_target{-1}, bad_w{-1};
    std::atomic<bool> canonical{true};
    parallel_for(n, T, [&](int64_t b, int64_t e, int) {
        bool canon = true;
        for (int64_t u = b; u < e; ++u) {
            const int32_t r0 = rp[u], len = rp[u + 1] - rp[u];
            const size_t toff = 4 + 4 * size_t(u + 1) + 4 * size_t(r0);
            if (len > 0) std::memcpy(&g->out_col[r0], tf.p + toff, 4 * size_t(len));
            for (int32_t k = 0; k < len; ++k) {
                const int32_t v = g->out_col[r0 + k];
                if (v < 0 || v >= n) bad_target.store(static_cast<int>(u));
                if (k > 0 && v <= g->out_col[r0 + k - 1]) canon = false;
            }
            if (have_w) {
                const size_t woff = 4 + 4 * size_t(u + 1) + 8 * size_t(r0);
                if (wf.i32(woff - 4) != len) bad_w.store(static_cast<int>(u));
                else if (len > 0) std::memcpy(&g->out_w[r0], wf.p + woff, 8 * size_t(len));
            } else {
                for (int32_t k = 0; k < len; ++k) g->out_w[r0 + k] = 1.0;
            }
        }
        if (!canon) canonical.store(false);
    });
    if (bad_target.load() >= 0)
        return set_error(GHET_ERR_IO, "Unable to read file %s (row %d has a target outside [0,%d))", path,
                         bad_target.load(), n);
    if (bad_w.load() >= 0)
        return set_error(GHET_ERR_IO, "Unable to read file %s.w (row %d length differs)", path, bad_w.load());
    if (!canonical.load()) {
        const int st = rebuild_rows(g);
        if (st != GHET_OK) return st;
    }
    *out = guard.release();
    return GHET_OK;
}

int ghet_host_graph_from_csr(int32_t n, const int32_t* rowptr, const int32_t* col, const double* w, int threads,
                             ghet_host_graph** out) {
    clear_error();
    if (!out) return set_error(GHET_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (n < 0 || !rowptr) return set_error(GHET_ERR_INVALID, "bad CSR");
    if (rowptr[0] != 0) return set_error(GHET_ERR_INVALID, "rowptr[0] = %d, expected 0", rowptr[0]);
    for (int64_t u = 0; u < n; ++u)
        if (rowptr[u + 1] < rowptr[u]) return set_error(GHET_ERR_INVALID, "offsets not monotone at vertex %lld", (long long)u);
    const int64_t m = rowptr[n];
    if (m > 0 && !col) return set_error(GHET_ERR_INVALID, "col is NULL");
    auto* g = new (std::nothrow) ghet_host_graph();
    if (!g) return set_error(GHET_ERR_NOMEM, "host allocation failed");
    std::unique_ptr<ghet_host_graph> guard(g);
    g->n = n;
    g->threads = hw_threads(threads);
    g->out_rowptr.assign(rowptr, rowptr + n + 1);
    g->out_col.assign(col, col + m);
    if (w) g->out_w.assign(w, w + m);
    else g->out_w.assign(static_cast<size_t>(m), 1.0);
    std::atomic<int64_t> bad{-1};
    std::atomic<bool> canonical{true};
    parallel_for(n, g->threads, [&](int64_t b, int64_t e, int) {
        bool canon = true;
        for (int64_t u = b; u < e; ++u)
            for (int32_t k = rowptr[u]; k < rowptr[u + 1]; ++k) {
                if (col[k] < 0 || col[k] >= n) bad.store(k);
                if (k > rowptr[u] && col[k] <= col[k - 1]) canon = false;
            }
        if (!canon) canonical.store(false);
    });
    if (bad.load() >= 0) return set_error(GHET_ERR_INVALID, "column index out of range at edge %lld", (long long)bad.load());
    if (!canonical.load()) {
        const int st = rebuild_rows(g);
        if (st != GHET_OK) return st;
    }
    *out = guard.release();
    return GHET_OK;
}

}  // extern "C"
