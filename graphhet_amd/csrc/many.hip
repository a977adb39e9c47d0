// Many small graphs in one captured launch sequence, and the autotuner.
//
// No reference counterpart (the reference handles one graph per call).
// Packed graphs become one disjoint union: graph i's local ids are offset by
// voff[i].  WCC and TC of a disjoint union are exactly the per-graph results
// (components and triangles never span graphs), so one pass over the union
// replaces `count` separate calls; WCC labels (smallest id of the component)
// are mapped back to local ids.  The device work of a call is captured into a
// hipGraph and launched once; it contains no memset nodes (fills are kernels).
//
// The autotuner times candidate launch shapes on the caller's own input and
// keeps the fastest; outputs are compared across candidates and must agree
// bit for bit (the shapes change speed only).
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>

#include "graph_internal.hpp"

namespace ghet {
int sssp_base_delta(ghet_graph* g, double* out, bool size_term);
namespace {

__global__ void k_localize(int32_t* __restrict__ lab, int32_t n, const int32_t* __restrict__ voff, int32_t count) {
    for (int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < n;
         v += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        int32_t lo = 0, hi = count;  // largest i with voff[i] <= v
        while (hi - lo > 1) {
            const int32_t mid = (lo + hi) >> 1;
            if (voff[mid] <= v) lo = mid;
            else hi = mid;
        }
        lab[v] -= voff[lo];
    }
}

// Process-wide cache of tuned shapes, keyed by a graph signature.
std::mutex g_tune_mu;
std::map<std::string, Tuning> g_tune_cache;

std::string signature(const char* what, int64_t n, int64_t m) {
    std::ostringstream os;
    os << what << ":" << n << ":" << m;
    return os.str();
}

bool cache_lookup(const std::string& key, Tuning* t) {
    std::lock_guard<std::mutex> l(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it == g_tune_cache.end()) return false;
    *t = it->second;
    return true;
}
void cache_store(const std::string& key, const Tuning& t) {
    std::lock_guard<std::mutex> l(g_tune_mu);
    g_tune_cache[key] = t;
}

// Builds the disjoint union of packed graphs as one device graph (out-CSR only).
int build_union(int32_t count, const int32_t* voff, const int32_t* rowptr, const int32_t* col, ghet_graph** out,
                int32_t** d_voff) {
    if (count < 0 || !voff || (count > 0 && (!rowptr || !col)))
        return set_error(GHET_ERR_INVALID, "bad packed-graph arguments");
    if (voff[0] != 0) return set_error(GHET_ERR_INVALID, "voff[0] must be 0");
    for (int32_t i = 0; i < count; ++i)
        if (voff[i + 1] < voff[i]) return set_error(GHET_ERR_INVALID, "voff not monotone at graph %d", i);
    const int32_t n = voff[count];
    if (rowptr[0] != 0) return set_error(GHET_ERR_INVALID, "rowptr[0] must be 0");
    const int64_t m = rowptr[n];
    std::vector<int32_t> gcol(static_cast<size_t>(m));
    int bad = -1;
    parallel_for(count, hw_threads(0), [&](int64_t b, int64_t e, int) {
        for (int64_t i = b; i < e; ++i) {
            const int32_t lo = voff[i], sz = voff[i + 1] - voff[i];
            for (int64_t k = rowptr[lo]; k < rowptr[voff[i + 1]]; ++k) {
                if (col[k] < 0 || col[k] >= sz) bad = static_cast<int>(i);
                gcol[k] = col[k] + lo;
            }
        }
    });
    if (bad >= 0) return set_error(GHET_ERR_INVALID, "graph %d has a column id outside its vertex range", bad);
    GHET_CHECK(ghet_graph_create(n, rowptr, gcol.data(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                 GHET_GRAPH_NO_IN | GHET_GRAPH_NO_WEIGHTS, out));
    void* p;
    GHET_CHECK((*out)->ws.get(WS_MISC, 4 * (size_t(count) + 1), &p));
    *d_voff = static_cast<int32_t*>(p);
    GHET_HIP(hipMemcpy(*d_voff, voff, 4 * (size_t(count) + 1), hipMemcpyHostToDevice));
    return GHET_OK;
}

struct GraphExec {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~GraphExec() {
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
    }
};

// Captures `enqueue(stream)` into a hipGraph and instantiates it.
template <typename F>
int capture(hipStream_t s, GraphExec& ge, F&& enqueue) {
    GHET_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int st = enqueue(s);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &graph);
    if (st != GHET_OK) {
        if (graph) hipGraphDestroy(graph);
        return st;
    }
    if (e != hipSuccess) return set_error(GHET_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    ge.graph = graph;
    GHET_HIP(hipGraphInstantiate(&ge.exec, graph, nullptr, nullptr, 0));
    return GHET_OK;
}

// Times `run(tuning)` for each candidate; keeps the fastest in g->tune.
template <typename Apply, typename Run>
int tune_over(ghet_graph* g, const std::vector<Tuning>& cands, Apply&& apply, Run&& run, hipStream_t s,
              std::vector<int64_t>* reference_out) {
    double best = 1e300;
    Tuning best_t = g->tune;
    std::vector<int64_t> first;
    for (size_t c = 0; c < cands.size(); ++c) {
        Tuning t = g->tune;
        apply(t, cands[c]);
        std::vector<int64_t> outv;
        GHET_HIP(hipEventRecord(g->ev_k0, s));
        GHET_CHECK(run(t, &outv));
        GHET_HIP(hipEventRecord(g->ev_k1, s));
        GHET_HIP(hipEventSynchronize(g->ev_k1));
        float ms = 0;
        GHET_HIP(hipEventElapsedTime(&ms, g->ev_k0, g->ev_k1));
        if (c == 0) first = outv;
        else if (outv != first)
            return set_error(GHET_ERR_HIP, "autotune: candidate %zu changed the result (must never happen)", c);
        if (ms < best) {
            best = ms;
            best_t = t;
        }
    }
    g->tune = best_t;
    if (reference_out) *reference_out = first;
    return GHET_OK;
}

std::vector<Tuning> wcc_candidates() {
    std::vector<Tuning> v;
    for (int light : {8, 16, 32, 64})
        for (int w : {32, 64}) {
            Tuning t;
            t.wcc_light = light;
            t.wcc_waves_per_cu = w;
            v.push_back(t);
        }
    return v;
}
std::vector<Tuning> tc_candidates() {
    std::vector<Tuning> v;
    for (int ept : {2, 4, 8, 16, 32})
        for (int w : {16, 32, 64}) {
            Tuning t;
            t.tc_ept = ept;
            t.tc_waves_per_cu = w;
            v.push_back(t);
        }
    return v;
}

int tune_wcc(ghet_graph* g, int32_t* parent, hipStream_t s) {
    return tune_over(
        g, wcc_candidates(), [](Tuning& t, const Tuning& c) { t.wcc_light = c.wcc_light; t.wcc_waves_per_cu = c.wcc_waves_per_cu; },
        [&](const Tuning& t, std::vector<int64_t>* out) -> int {
            GHET_CHECK(enqueue_wcc_plain(g, parent, t.wcc_light, t.wcc_waves_per_cu, s));
            std::vector<int32_t> h(static_cast<size_t>(g->n));
            GHET_HIP(hipMemcpyAsync(h.data(), parent, 4 * h.size(), hipMemcpyDeviceToHost, s));
            GHET_HIP(hipStreamSynchronize(s));
            out->assign(h.begin(), h.end());
            return GHET_OK;
        },
        s, nullptr);
}

int tune_tc(ghet_graph* g, unsigned long long* cnt, hipStream_t s) {
    return tune_over(
        g, tc_candidates(), [](Tuning& t, const Tuning& c) { t.tc_ept = c.tc_ept; t.tc_waves_per_cu = c.tc_waves_per_cu; },
        [&](const Tuning& t, std::vector<int64_t>* out) -> int {
            GHET_CHECK(enqueue_tc(g, cnt, t.tc_ept, t.tc_waves_per_cu, s));
            out->resize(static_cast<size_t>(g->n));
            GHET_HIP(hipMemcpyAsync(out->data(), cnt, 8 * out->size(), hipMemcpyDeviceToHost, s));
            GHET_HIP(hipStreamSynchronize(s));
            return GHET_OK;
        },
        s, nullptr);
}

bool want_tuning(const ghet_opts* o) { return !(o && o->size >= sizeof(ghet_opts) && (o->reserved[1] & 1)); }

}  // namespace
}  // namespace ghet

using namespace ghet;

extern "C" {

int ghet_wcc_many(int32_t count, const int32_t* voff, const int32_t* rowptr, const int32_t* col, int32_t* label_out,
                  const ghet_opts* opts, ghet_run_stats* stats) {
    clear_error();
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (count == 0) return GHET_OK;
    ghet_graph* g = nullptr;
    int32_t* d_voff = nullptr;
    int st = build_union(count, voff, rowptr, col, &g, &d_voff);
    if (st != GHET_OK) {
        ghet_graph_destroy(g);
        return st;
    }
    auto run = [&]() -> int {
        hipStream_t s = g->stream;
        const int32_t n = g->n;
        void* p;
        GHET_CHECK(g->ws.get(WS_PARENT, 4 * size_t(std::max(n, 1)), &p));
        int32_t* parent = static_cast<int32_t*>(p);
        const std::string key = signature("wcc", n, g->out.m);
        if (!cache_lookup(key, &g->tune) && want_tuning(opts)) {
            GHET_CHECK(tune_wcc(g, parent, s));
            cache_store(key, g->tune);
        }
        GraphExec ge;
        GHET_CHECK(capture(s, ge, [&](hipStream_t cs) -> int {
            GHET_CHECK(enqueue_wcc_plain(g, parent, g->tune.wcc_light, g->tune.wcc_waves_per_cu, cs));
            hipLaunchKernelGGL(k_localize, dim3(static_cast<unsigned>(std::min<int64_t>(ceil_div(n, 256), 8192))),
                               dim3(256), 0, cs, parent, n, d_voff, count);
            GHET_HIP(hipGetLastError());
            return GHET_OK;
        }));
        GHET_HIP(hipEventRecord(g->ev_begin, s));
        GHET_HIP(hipGraphLaunch(ge.exec, s));
        GHET_HIP(hipEventRecord(g->ev_end, s));
        if (label_out) GHET_HIP(hipMemcpyAsync(label_out, parent, 4 * size_t(n), hipMemcpyDeviceToHost, s));
        GHET_HIP(hipStreamSynchronize(s));
        if (stats) {
            float ms = 0;
            GHET_HIP(hipEventElapsedTime(&ms, g->ev_begin, g->ev_end));
            stats->device_ms = ms;
            stats->levels = 1;
            stats->extra[0] = g->tune.wcc_light;
            stats->extra[1] = g->tune.wcc_waves_per_cu;
        }
        return GHET_OK;
    };
    st = run();
    ghet_graph_destroy(g);
    return st;
}

int ghet_tc_many(int32_t count, const int32_t* voff, const int32_t* rowptr, const int32_t* col, int64_t* count_out,
                 const ghet_opts* opts, ghet_run_stats* stats) {
    clear_error();
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (count == 0) return GHET_OK;
    ghet_graph* g = nullptr;
    int32_t* d_voff = nullptr;
    int st = build_union(count, voff, rowptr, col, &g, &d_voff);
    if (st != GHET_OK) {
        ghet_graph_destroy(g);
        return st;
    }
    auto run = [&]() -> int {
        hipStream_t s = g->stream;
        const int32_t n = g->n;
        GHET_CHECK(tc_prepare(g, s));
        void* p;
        GHET_CHECK(g->ws.get(WS_TC_COUNT, 8 * size_t(std::max(n, 1)), &p));
        auto* cnt = static_cast<unsigned long long*>(p);
        const std::string key = signature(g->tc.symmetric ? "tcs" : "tcg", n, g->out.m);
        if (!cache_lookup(key, &g->tune) && want_tuning(opts)) {
            GHET_CHECK(tune_tc(g, cnt, s));
            cache_store(key, g->tune);
        }
        GraphExec ge;
        GHET_CHECK(capture(s, ge, [&](hipStream_t cs) -> int {
            return enqueue_tc(g, cnt, g->tune.tc_ept, g->tune.tc_waves_per_cu, cs);
        }));
        GHET_HIP(hipEventRecord(g->ev_begin, s));
        GHET_HIP(hipGraphLaunch(ge.exec, s));
        GHET_HIP(hipEventRecord(g->ev_end, s));
        if (count_out) GHET_HIP(hipMemcpyAsync(count_out, cnt, 8 * size_t(n), hipMemcpyDeviceToHost, s));
        GHET_HIP(hipStreamSynchronize(s));
        if (stats) {
            float ms = 0;
            GHET_HIP(hipEventElapsedTime(&ms, g->ev_begin, g->ev_end));
            stats->device_ms = ms;
            stats->levels = 1;
            stats->extra[0] = g->tune.tc_ept;
            stats->extra[1] = g->tune.tc_waves_per_cu;
            stats->extra[2] = g->tc.symmetric ? 1.0 : 0.0;
        }
        return GHET_OK;
    };
    st = run();
    ghet_graph_destroy(g);
    return st;
}

// Tunes TC shapes and the SSSP bucket width on g; results go to g's tuning
// (used by later calls on g) and, when cache_path is given, are appended to
// that text file as "<signature> <field>=<value> ...".  A cache entry for the
// same signature is used instead of timing.
int ghet_autotune(ghet_graph* g, const char* cache_path, int32_t* best_block_out) {
    clear_error();
    if (!g) return set_error(GHET_ERR_INVALID, "graph handle is NULL");
    if (hipSetDevice(g->device) != hipSuccess) return set_error(GHET_ERR_HIP, "hipSetDevice failed");
    hipStream_t s = g->stream;
    const std::string key = signature("graph", g->n, g->out.m);
    if (cache_path) {
        std::ifstream in(cache_path);
        std::string line;
        while (std::getline(in, line)) {
            std::istringstream is(line);
            std::string k;
            is >> k;
            if (k != key) continue;
            std::string kv;
            Tuning t = g->tune;
            while (is >> kv) {
                const size_t eq = kv.find('=');
                if (eq == std::string::npos) continue;
                const std::string f = kv.substr(0, eq);
                const int val = std::atoi(kv.c_str() + eq + 1);
                if (f == "tc_ept" && val > 0) t.tc_ept = val;
                else if (f == "tc_waves" && val > 0) t.tc_waves_per_cu = val;
                else if (f == "sssp_class" && val >= 0 && val < 4) t.sssp_class = val;
            }
            t.tuned = true;
            g->tune = t;
        }
    }
    if (!g->tune.tuned) {
        if (g->out.present() && g->n > 0) {
            GHET_CHECK(tc_prepare(g, s));
            void* p;
            GHET_CHECK(g->ws.get(WS_TC_COUNT, 8 * size_t(g->n), &p));
            GHET_CHECK(tune_tc(g, static_cast<unsigned long long*>(p), s));
        }
        if (g->out.w && g->n > 0) {
            // SSSP bucket width: time the candidate classes from the max-degree vertex
            int32_t src = 0;
            for (int32_t v = 1; v < g->n; ++v)
                if (g->h_out_rowptr[v + 1] - g->h_out_rowptr[v] > g->h_out_rowptr[src + 1] - g->h_out_rowptr[src]) src = v;
            double best = 1e300;
            g->tune.sssp_class = 1;
            double base0 = 0;
            GHET_CHECK(sssp_base_delta(g, &base0, true));
            for (int cls = 0; cls < 4; ++cls) {
                ghet_opts o{};
                o.size = sizeof(o);
                o.sssp_delta = base0 * std::ldexp(1.0, cls - 1);
                ghet_run_stats stt{};
                GHET_CHECK(run_sssp(g, src, 1.7976931348623157e308, nullptr, nullptr, &o, &stt));
                if (stt.device_ms < best) {
                    best = stt.device_ms;
                    g->tune.sssp_class = cls;
                }
            }
        }
        g->tune.tuned = true;
        if (cache_path) {
            std::ofstream out(cache_path, std::ios::app);
            out << key << " tc_ept=" << g->tune.tc_ept << " tc_waves=" << g->tune.tc_waves_per_cu
                << " sssp_class=" << g->tune.sssp_class << "\n";
        }
    }
    if (best_block_out) *best_block_out = g->tune.tc_ept;
    return GHET_OK;
}

}  // extern "C"
