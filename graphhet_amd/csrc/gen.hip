// On-GPU R-MAT generation: the graph ghet_gen_rmat builds on the host, built
// directly in device memory as a ready ghet_graph (out-, in- and all-CSR with
// weights), plus ghet_graph_download to bring any CSR back to the host.
//
// No reference counterpart on the device: the reference builds inputs on the
// host with one sorted insert per edge (CPU_adj_list.hpp:110-122) and copies
// them with toCSR (GPU_csr.hpp:282-339).  Here the whole pipeline stays on the
// GPU:
//   1. one thread per sample k draws its `scale` quadrant choices from the
//      counter-based SplitMix64 stream (the host generator's draws, bit for
//      bit) and writes the key u * 2^scale + v (self-loops: 2^(2 scale),
//      sorting last) with value k;
//   2. a stable LSD radix sort over the 2 scale + 1 key bits orders samples by
//      (u, v), equal pairs by k;
//   3. the last sample of every run of equal keys is kept (the latest sample's
//      weight wins, as on the host), compacted by a scan: out_col = v, the
//      weight from sample k's weight draw, the row ids u;
//   4. out-offsets by binary search of the sorted rows;
//   5. in-CSR: a stable sort of the targets with edge ids as values gives each
//      in-row its sources in ascending u; offsets by binary search again;
//   6. all-CSR: per vertex its in-list then its out-list.
// Temporary memory peaks near 50 bytes per sample (about 13 GB at scale 24).
#include <cmath>
#include <cstring>

#include "graph_internal.hpp"

#include <rocprim/rocprim.hpp>

namespace ghet {

int graph_alloc_arena(int32_t n, int64_t m_out, bool want_in, int64_t m_in, bool want_all, int64_t m_all,
                      bool want_w, ghet_graph** out);
void graph_free(ghet_graph* g);

namespace {

// SplitMix64 (Steele, Lea, Flood 2014), the same counter-based stream as the
// host generator: draw(key, j) = mix64(key + (j+1) * golden), key =
// mix64(seed + (k+1) * golden).  The GPU tests compare every CSR of the
// generated graph with the host generator's, bit for bit.
constexpr uint64_t kGoldenD = 0x9E3779B97F4A7C15ULL;
__device__ __forceinline__ uint64_t mix64_d(uint64_t x) {
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ULL;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBULL;
    x ^= x >> 31;
    return x;
}
__device__ __forceinline__ uint64_t stream_key_d(uint64_t seed, uint64_t k) { return mix64_d(seed + kGoldenD * (k + 1)); }
__device__ __forceinline__ uint64_t draw_d(uint64_t key, uint64_t j) { return mix64_d(key + kGoldenD * (j + 1)); }
__device__ __forceinline__ double weight_from_d(uint64_t r) {
    return static_cast<double>((r >> 11) + 1) * 0x1.0p-53;
}

__global__ void __launch_bounds__(256) k_rmat_keys(int64_t S, int scale, uint64_t t1, uint64_t t2, uint64_t t3,
                                                   uint64_t seed, uint64_t* __restrict__ keys,
                                                   uint32_t* __restrict__ vals) {
    const uint64_t loop_key = 1ULL << (2 * scale);
    for (int64_t k = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; k < S;
         k += static_cast<int64_t>(gridDim.x) * 256) {
        const uint64_t key = stream_key_d(seed, static_cast<uint64_t>(k));
        uint32_t u = 0, v = 0;
        uint64_t r = 0;
        for (int j = 0; j < scale; ++j) {
            if ((j & 1) == 0) r = draw_d(key, static_cast<uint64_t>(j >> 1));
            const uint64_t x = (j & 1) ? (r >> 32) : (r & 0xFFFFFFFFULL);
            const uint32_t q = x < t1 ? 0u : x < t2 ? 1u : x < t3 ? 2u : 3u;
            u = (u << 1) | (q >> 1);
            v = (v << 1) | (q & 1u);
        }
        keys[k] = (u == v) ? loop_key : ((static_cast<uint64_t>(u) << scale) | v);
        vals[k] = static_cast<uint32_t>(k);
    }
}

// flag[i] = 1 for the last sample of each run of equal (non-loop) keys.
__global__ void __launch_bounds__(256) k_last_of_run(const uint64_t* __restrict__ keys, int64_t S, uint64_t loop_key,
                                                     int32_t* __restrict__ flag) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < S;
         i += static_cast<int64_t>(gridDim.x) * 256) {
        const uint64_t k = keys[i];
        flag[i] = (k != loop_key && (i + 1 == S || keys[i + 1] != k)) ? 1 : 0;
    }
}

__global__ void __launch_bounds__(256) k_compact(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                                 const int32_t* __restrict__ flag, const int32_t* __restrict__ pos,
                                                 int64_t S, int scale, uint64_t seed, uint64_t wdraw,
                                                 int32_t* __restrict__ row, int32_t* __restrict__ col,
                                                 double* __restrict__ w) {
    const uint64_t vmask = (1ULL << scale) - 1;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < S;
         i += static_cast<int64_t>(gridDim.x) * 256) {
        if (!flag[i]) continue;
        const int32_t j = pos[i];
        const uint64_t k = keys[i];
        row[j] = static_cast<int32_t>(k >> scale);
        col[j] = static_cast<int32_t>(k & vmask);
        if (w) w[j] = weight_from_d(draw_d(stream_key_d(seed, vals[i]), wdraw));
    }
}

// rowptr[r] = first index of sorted ids with id >= r, r in [0, n].
__global__ void __launch_bounds__(256) k_offsets(const int32_t* __restrict__ ids, int64_t m, int32_t n,
                                                 int32_t* __restrict__ rowptr) {
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; r <= n;
         r += static_cast<int64_t>(gridDim.x) * 256) {
        int64_t lo = 0, hi = m;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (ids[mid] < r) lo = mid + 1;
            else hi = mid;
        }
        rowptr[r] = static_cast<int32_t>(lo);
    }
}

__global__ void __launch_bounds__(256) k_iota(uint32_t* __restrict__ x, int64_t m) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < m;
         i += static_cast<int64_t>(gridDim.x) * 256)
        x[i] = static_cast<uint32_t>(i);
}

// In-CSR entries from the target-sorted edge ids: source u and the weight of
// edge u -> v.
__global__ void __launch_bounds__(256) k_in_fill(const uint32_t* __restrict__ perm, int64_t m,
                                                 const int32_t* __restrict__ row, const double* __restrict__ w,
                                                 int32_t* __restrict__ in_col, double* __restrict__ in_w) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < m;
         i += static_cast<int64_t>(gridDim.x) * 256) {
        const uint32_t e = perm[i];
        in_col[i] = row[e];
        if (in_w) in_w[i] = w[e];
    }
}

__global__ void __launch_bounds__(256) k_all_rowptr(const int32_t* __restrict__ orp, const int32_t* __restrict__ irp,
                                                    int32_t n, int32_t* __restrict__ arp) {
    for (int64_t v = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; v <= n;
         v += static_cast<int64_t>(gridDim.x) * 256)
        arp[v] = orp[v] + irp[v];
}

// One wave per vertex: its in-list, then its out-list.
__global__ void __launch_bounds__(256) k_all_fill(const int32_t* __restrict__ orp, const int32_t* __restrict__ ocol,
                                                  const int32_t* __restrict__ irp, const int32_t* __restrict__ icol,
                                                  const int32_t* __restrict__ arp, int32_t n,
                                                  int32_t* __restrict__ acol) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = (static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x) >> 6;
    const int64_t waves = (static_cast<int64_t>(gridDim.x) * 256) >> 6;
    for (int64_t v = wave; v < n; v += waves) {
        const int32_t d = arp[v], ib = irp[v], ie = irp[v + 1], ob = orp[v], oe = orp[v + 1];
        for (int32_t i = ib + lane; i < ie; i += 64) acol[d + (i - ib)] = icol[i];
        for (int32_t i = ob + lane; i < oe; i += 64) acol[d + (ie - ib) + (i - ob)] = ocol[i];
    }
}

unsigned grid_for(int64_t items) {
    return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(ceil_div(items, 256), 65536)));
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) hipFree(p);
    }
    int alloc(size_t bytes) {
        hipError_t e = hipMalloc(&p, std::max<size_t>(bytes, 256));
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            return set_error(GHET_ERR_NOMEM, "generator: hipMalloc(%zu) failed", bytes);
        }
        return GHET_OK;
    }
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

}  // namespace

// all-CSR (per vertex its in-list, then its out-list) from device out- and
// in-CSRs, on stream s.
void launch_all_csr(const int32_t* orp, const int32_t* ocol, const int32_t* irp, const int32_t* icol, int32_t n,
                    int32_t* arp, int32_t* acol, hipStream_t s) {
    hipLaunchKernelGGL(k_all_rowptr, dim3(grid_for(int64_t(n) + 1)), dim3(256), 0, s, orp, irp, n, arp);
    hipLaunchKernelGGL(k_all_fill, dim3(grid_for(int64_t(n) * 64)), dim3(256), 0, s, orp, ocol, irp, icol, arp, n,
                       acol);
}

}  // namespace ghet

using namespace ghet;

extern "C" {

int ghet_graph_gen_rmat(int scale, double edge_factor, double a, double b, double c, uint64_t seed, uint32_t flags,
                        ghet_graph** out) {
    clear_error();
    if (!out) return set_error(GHET_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (scale < 1 || scale > 30) return set_error(GHET_ERR_INVALID, "scale %d outside [1,30]", scale);
    if (!(edge_factor > 0) || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        return set_error(GHET_ERR_INVALID, "bad R-MAT parameters");
    if (flags & GHET_GRAPH_UNDIRECTED) return set_error(GHET_ERR_INVALID, "R-MAT graphs are directed");
    const double samples_d = edge_factor * std::ldexp(1.0, scale);
    if (samples_d >= 2147483647.0) return set_error(GHET_ERR_UNSUPPORTED, "too many samples (%g)", samples_d);
    if (ghet_device_count() < 1) return set_error(GHET_ERR_NODEVICE, "no GPU visible to this process");
    const int64_t S = static_cast<int64_t>(samples_d);
    const int32_t n = static_cast<int32_t>(1LL << scale);
    const bool want_w = !(flags & GHET_GRAPH_NO_WEIGHTS);
    const bool want_in = !(flags & GHET_GRAPH_NO_IN);
    const bool want_all = want_in && !(flags & GHET_GRAPH_NO_ALL);
    // the host generator's quadrant thresholds and weight draw index
    const uint64_t t1 = static_cast<uint64_t>(a * 4294967296.0);
    const uint64_t t2 = static_cast<uint64_t>((a + b) * 4294967296.0);
    const uint64_t t3 = static_cast<uint64_t>((a + b + c) * 4294967296.0);
    const uint64_t wdraw = static_cast<uint64_t>((scale + 1) / 2);
    const uint64_t loop_key = 1ULL << (2 * scale);
    // Temporaries are declared before the stream guard, so every exit path
    // drains the stream before freeing them.
    DevBuf kA, kB, vA, vB, tmp, tmp2, flag, pos, aux, rows;
    hipStream_t s;
    GHET_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() {
            hipStreamSynchronize(s);
            hipStreamDestroy(s);
        }
    } sg{s};

    // 1-2. keyed samples, stable radix sort by (u, v)
    GHET_CHECK(kA.alloc(8 * size_t(S)));
    GHET_CHECK(kB.alloc(8 * size_t(S)));
    GHET_CHECK(vA.alloc(4 * size_t(S)));
    GHET_CHECK(vB.alloc(4 * size_t(S)));
    hipLaunchKernelGGL(k_rmat_keys, dim3(grid_for(S)), dim3(256), 0, s, S, scale, t1, t2, t3, seed,
                       kA.as<uint64_t>(), vA.as<uint32_t>());
    GHET_HIP(hipGetLastError());
    const unsigned key_bits = static_cast<unsigned>(2 * scale + 1);
    {
        rocprim::double_buffer<uint64_t> kd(kA.as<uint64_t>(), kB.as<uint64_t>());
        rocprim::double_buffer<uint32_t> vd(vA.as<uint32_t>(), vB.as<uint32_t>());
        size_t tb = 0;
        GHET_HIP(rocprim::radix_sort_pairs(nullptr, tb, kd, vd, S, 0, key_bits, s));
        GHET_CHECK(tmp.alloc(tb));
        GHET_HIP(rocprim::radix_sort_pairs(tmp.p, tb, kd, vd, S, 0, key_bits, s));
        if (kd.current() != kA.as<uint64_t>()) std::swap(kA.p, kB.p);
        if (vd.current() != vA.as<uint32_t>()) std::swap(vA.p, vB.p);
    }
    // 3. keep the last sample of each (u, v) run
    GHET_CHECK(flag.alloc(4 * size_t(S)));
    GHET_CHECK(pos.alloc(4 * (size_t(S) + 1)));
    GHET_CHECK(aux.alloc(4 * size_t(scan_aux_ints(S))));
    hipLaunchKernelGGL(k_last_of_run, dim3(grid_for(S)), dim3(256), 0, s, kA.as<uint64_t>(), S, loop_key,
                       flag.as<int32_t>());
    GHET_CHECK(exclusive_scan_i32(flag.as<int32_t>(), pos.as<int32_t>(), S, aux.as<int32_t>(), s));
    int32_t m32 = 0;
    GHET_HIP(hipMemcpyAsync(&m32, pos.as<int32_t>() + S, 4, hipMemcpyDeviceToHost, s));
    GHET_HIP(hipStreamSynchronize(s));
    const int64_t m = m32;
    const int64_t m_all = want_all ? 2 * m : 0;
    if (m_all > INT32_MAX) return set_error(GHET_ERR_UNSUPPORTED, "all-CSR has %lld entries (> 2^31-1)", (long long)m_all);

    ghet_graph* g = nullptr;
    GHET_CHECK(graph_alloc_arena(n, m, want_in, m, want_all, m_all, want_w, &g));
    struct GraphGuard {
        ghet_graph* g;
        ~GraphGuard() {
            if (g) graph_free(g);
        }
    } gg{g};
    // row ids of the kept edges
    GHET_CHECK(rows.alloc(4 * size_t(m)));
    hipLaunchKernelGGL(k_compact, dim3(grid_for(S)), dim3(256), 0, s, kA.as<uint64_t>(), vA.as<uint32_t>(),
                       flag.as<int32_t>(), pos.as<int32_t>(), S, scale, seed, wdraw, rows.as<int32_t>(), g->out.col,
                       g->out.w);
    // 4. out-offsets
    hipLaunchKernelGGL(k_offsets, dim3(grid_for(int64_t(n) + 1)), dim3(256), 0, s, rows.as<int32_t>(), m, n,
                       g->out.rowptr);
    GHET_HIP(hipGetLastError());
    // 5. in-CSR: stable sort of targets carrying edge ids
    if (want_in) {
        uint32_t* tk = reinterpret_cast<uint32_t*>(kA.p);  // the sample keys are consumed
        uint32_t* tk2 = reinterpret_cast<uint32_t*>(kB.p);
        uint32_t* ev = vA.as<uint32_t>();
        uint32_t* ev2 = vB.as<uint32_t>();
        GHET_HIP(hipMemcpyAsync(tk, g->out.col, 4 * size_t(m), hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(k_iota, dim3(grid_for(m)), dim3(256), 0, s, ev, m);
        rocprim::double_buffer<uint32_t> kd(tk, tk2);
        rocprim::double_buffer<uint32_t> vd(ev, ev2);
        size_t tb = 0;
        GHET_HIP(rocprim::radix_sort_pairs(nullptr, tb, kd, vd, m, 0, static_cast<unsigned>(scale), s));
        GHET_CHECK(tmp2.alloc(tb));
        GHET_HIP(rocprim::radix_sort_pairs(tmp2.p, tb, kd, vd, m, 0, static_cast<unsigned>(scale), s));
        hipLaunchKernelGGL(k_in_fill, dim3(grid_for(m)), dim3(256), 0, s, vd.current(), m, rows.as<int32_t>(),
                           g->out.w, g->in.col, g->in.w);
        hipLaunchKernelGGL(k_offsets, dim3(grid_for(int64_t(n) + 1)), dim3(256), 0, s,
                           reinterpret_cast<const int32_t*>(kd.current()), m, n, g->in.rowptr);
        GHET_HIP(hipGetLastError());
    }
    // 6. all-CSR
    if (want_all) {
        launch_all_csr(g->out.rowptr, g->out.col, g->in.rowptr, g->in.col, n, g->all.rowptr, g->all.col, s);
        GHET_HIP(hipGetLastError());
    }
    // host copies of the offsets (the layout builders read them)
    g->h_out_rowptr.resize(size_t(n) + 1);
    GHET_HIP(hipMemcpyAsync(g->h_out_rowptr.data(), g->out.rowptr, 4 * (size_t(n) + 1), hipMemcpyDeviceToHost, s));
    if (want_in) {
        g->h_in_rowptr.resize(size_t(n) + 1);
        GHET_HIP(hipMemcpyAsync(g->h_in_rowptr.data(), g->in.rowptr, 4 * (size_t(n) + 1), hipMemcpyDeviceToHost, s));
    }
    if (want_all) {
        g->h_all_rowptr.resize(size_t(n) + 1);
        GHET_HIP(hipMemcpyAsync(g->h_all_rowptr.data(), g->all.rowptr, 4 * (size_t(n) + 1), hipMemcpyDeviceToHost, s));
    }
    GHET_HIP(hipStreamSynchronize(s));
    *out = g;
    gg.g = nullptr;
    return GHET_OK;
}

int ghet_graph_download(const ghet_graph* g, int kind, int32_t* rowptr, int32_t* col, double* w) {
    clear_error();
    if (!g) return set_error(GHET_ERR_INVALID, "graph handle is NULL");
    const DevCSR* c = nullptr;
    switch (kind) {
        case GHET_VIEW_OUT: c = &g->out; break;
        case GHET_VIEW_IN: c = &g->in; break;
        case GHET_VIEW_ALL: c = &g->all; break;
        default: return set_error(GHET_ERR_INVALID, "download: view kind %d not held on the device", kind);
    }
    if (!c->present()) return set_error(GHET_ERR_INVALID, "download: the graph has no such CSR");
    if (w && !c->w) return set_error(GHET_ERR_INVALID, "download: the CSR has no weights");
    GHET_HIP(hipSetDevice(g->device));
    if (rowptr) GHET_HIP(hipMemcpy(rowptr, c->rowptr, 4 * (size_t(g->n) + 1), hipMemcpyDeviceToHost));
    if (col && c->m > 0) GHET_HIP(hipMemcpy(col, c->col, 4 * size_t(c->m), hipMemcpyDeviceToHost));
    if (w && c->m > 0) GHET_HIP(hipMemcpy(w, c->w, 8 * size_t(c->m), hipMemcpyDeviceToHost));
    return GHET_OK;
}

}  // extern "C"
