// Single-source shortest paths in FP64 on gfx950: near-far bucketing.
//
// Reference: csr_sssp (GPU_SSSP_csr.cuh:253-319) is frontier Bellman-Ford with
// an 8-byte (u<<32|e) candidate queue, two launches and five syncs per round;
// semantics of CPU_shortest_paths (CPU_shortest_paths.hpp:19-64), Dijkstra with
// relaxation d[u]+w < d[v] and DBL_MAX for unreached.
//
// Result identity.  Weights are >= 0 and FP64 addition is monotone, so every
// label-correcting schedule and Dijkstra converge to the same fixpoint: the
// minimum over walks of the left-to-right rounded path sum.  The bucket width
// therefore changes speed, never a bit of the output.
//
// Each round expands the near queue edge-balanced (degree scan + 2048-edge
// tiles); an improved vertex goes to the next near queue when its new distance
// is below the threshold, else to the far list.  Per-vertex round/generation
// stamps keep both lists duplicate-free.  When the near queue drains, the far
// list is split at threshold + delta.
#include <cfloat>
#include <cmath>
#include <cstring>

#include "graph_internal.hpp"
#include "sched.hpp"

namespace ghet {
namespace {

struct SsspCtr {
    int32_t near_count;
    int32_t far_count[2];
    int32_t pad;
    unsigned long long near_edges;
    unsigned long long far_min_bits;
};

constexpr int kBT = 256;
constexpr int kItems = 8;
constexpr int kTile = kBT * kItems;

__device__ __forceinline__ int upper_bound_i32(const int32_t* a, int len, int32_t key) {
    int lo = 0, hi = len;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (a[mid] <= key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ void k_sssp_seed(double* dist, int32_t* queue, int32_t* stamp_near, int32_t src) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        dist[src] = 0.0;
        queue[0] = src;
        stamp_near[src] = 0;
    }
}

__global__ void k_round_reset(SsspCtr* c) {
    if (threadIdx.x == 0) {
        c->near_count = 0;
        c->near_edges = 0;
    }
}
__global__ void k_split_reset(SsspCtr* c, int other) {
    if (threadIdx.x == 0) {
        c->near_count = 0;
        c->near_edges = 0;
        c->far_count[other] = 0;
        c->far_min_bits = 0x7FF0000000000000ULL;  // +inf
    }
}

__global__ void k_frontier_deg(const int32_t* __restrict__ q, int32_t F, const int32_t* __restrict__ rowptr,
                               int32_t* __restrict__ deg) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < F) {
        int32_t u = q[i];
        deg[i] = rowptr[u + 1] - rowptr[u];
    }
}

__global__ void __launch_bounds__(kBT) k_relax(const int32_t* __restrict__ q, int32_t F,
                                               const int32_t* __restrict__ scan, int32_t E,
                                               const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                               const double* __restrict__ wts, double* dist, double threshold,
                                               int32_t* stamp_near, int32_t next_round, int32_t* stamp_far,
                                               int32_t far_gen, int32_t* __restrict__ next_q, int32_t* far_list,
                                               int far_sel, SsspCtr* ctr) {
    __shared__ int32_t s_scan[kTile];
    __shared__ int32_t s_start[kTile];
    __shared__ int32_t s_src[kTile];
    __shared__ int32_t s_out[kTile];
    __shared__ int32_t s_i0, s_span, s_cnt, s_base;
    __shared__ unsigned long long s_red[kBT / 64];
    const int tid = threadIdx.x;
    for (int64_t tile = blockIdx.x; tile * kTile < E; tile += gridDim.x) {
        const int32_t t0 = static_cast<int32_t>(tile * kTile);
        const int32_t t1 = static_cast<int32_t>(std::min<int64_t>(static_cast<int64_t>(t0) + kTile, E));
        if (tid == 0) {
            int i0 = upper_bound_i32(scan, F + 1, t0) - 1;
            int i1 = upper_bound_i32(scan, F + 1, t1 - 1) - 1;
            s_i0 = i0;
            s_span = i1 - i0 + 1;
            s_cnt = 0;
        }
        __syncthreads();
        const int i0 = s_i0, span = s_span;
        const bool use_lds = span <= kTile;
        if (use_lds) {
            for (int j = tid; j < span; j += kBT) {
                int32_t u = q[i0 + j];
                s_scan[j] = scan[i0 + j];
                s_start[j] = rowptr[u];
                s_src[j] = u;
            }
        }
        __syncthreads();
        for (int k = 0; k < kItems; ++k) {
            const int32_t e = t0 + k * kBT + tid;
            bool to_near = false;
            int32_t v = 0;
            if (e < t1) {
                int32_t pos, u;
                if (use_lds) {
                    int j = upper_bound_i32(s_scan, span, e) - 1;
                    pos = s_start[j] + (e - s_scan[j]);
                    u = s_src[j];
                } else {
                    int j = upper_bound_i32(scan + i0, span, e) - 1;
                    u = q[i0 + j];
                    pos = rowptr[u] + (e - scan[i0 + j]);
                }
                v = col[pos];
                const double nd = dist[u] + wts[pos];
                if (nd < dist[v]) {
                    const double old = atomic_min_nonneg(&dist[v], nd);
                    if (nd < old) {
                        if (nd < threshold) {
                            if (atomicExch(&stamp_near[v], next_round) != next_round) to_near = true;
                        } else if (atomicExch(&stamp_far[v], far_gen) != far_gen) {
                            int p = atomicAdd(&ctr->far_count[far_sel], 1);
                            far_list[p] = v;
                        }
                    }
                }
            }
            uint64_t mask = __ballot(to_near);
            if (mask) {
                int leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
                int base = 0;
                if (lane_id() == leader) base = atomicAdd(&s_cnt, __popcll(mask));
                base = __shfl(base, leader);
                if (to_near) s_out[base + lanes_below(mask)] = v;
            }
        }
        __syncthreads();
        if (tid == 0 && s_cnt) s_base = atomicAdd(&ctr->near_count, s_cnt);
        __syncthreads();
        unsigned long long od = 0;
        const int cnt = s_cnt, base = s_base;
        for (int j = tid; j < cnt; j += kBT) {
            int32_t v = s_out[j];
            next_q[base + j] = v;
            od += rowptr[v + 1] - rowptr[v];
        }
        od = wave_sum(od);
        if ((tid & 63) == 0) s_red[tid >> 6] = od;
        __syncthreads();
        if (tid == 0) {
            unsigned long long a = 0;
            for (int w = 0; w < kBT / 64; ++w) a += s_red[w];
            if (a) atomicAdd(&ctr->near_edges, a);
        }
        __syncthreads();
    }
}

// Splits far list `sel` at new_thr: entries below old_thr are stale (already
// settled at their current distance), [old_thr, new_thr) go to the near queue,
// the rest to far list 1-sel with the new generation stamp.
__global__ void k_far_split(const int32_t* __restrict__ far_in, int32_t count, const double* __restrict__ dist,
                            double old_thr, double new_thr, int32_t* stamp_near, int32_t next_round,
                            int32_t* stamp_far, int32_t new_gen, int32_t* __restrict__ near_q, int32_t* far_out,
                            int other, const int32_t* __restrict__ rowptr, SsspCtr* ctr) {
    unsigned long long od = 0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i - threadIdx.x < count;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        bool to_near = false, keep = false;
        int32_t v = 0;
        double d = 0;
        if (i < count) {
            v = far_in[i];
            d = dist[v];
            if (d >= old_thr) {
                if (d < new_thr) {
                    stamp_near[v] = next_round;
                    to_near = true;
                } else {
                    stamp_far[v] = new_gen;
                    keep = true;
                }
            }
        }
        int p = wave_append(&ctr->near_count, to_near);
        if (to_near) {
            near_q[p] = v;
            od += rowptr[v + 1] - rowptr[v];
        }
        int k = wave_append(&ctr->far_count[other], keep);
        if (keep) {
            far_out[k] = v;
            atomicMin(&ctr->far_min_bits, static_cast<unsigned long long>(__double_as_longlong(d)));
        }
    }
    od = wave_sum(od);
    if (lane_id() == 0 && od) atomicAdd(&ctr->near_edges, od);
}

// Parent = smallest u != v with an edge u->v and d[u] + w(u,v) == d[v]
// (atomicMin over out-edges: deterministic).  pre must hold INT_MAX
// beforehand; the source keeps INT_MAX (mapped to -1 afterwards).
__global__ void __launch_bounds__(256) k_sssp_parent(int32_t n, const int32_t* __restrict__ rowptr,
                                                     const int32_t* __restrict__ col, const double* __restrict__ w,
                                                     const double* __restrict__ dist, double inf, int32_t src,
                                                     int32_t* pre) {
    const int lane = lane_id();
    const int64_t groups = ceil_div(n, 64);
    const int64_t wave_global = (static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x) >> 6;
    const int64_t wave_stride = (static_cast<int64_t>(gridDim.x) * 256) >> 6;
    for (int64_t g = wave_global; g < groups; g += wave_stride) {
        const int64_t u = g * 64 + lane;
        int32_t b = 0, e = 0;
        double du = inf;
        if (u < n) {
            du = dist[u];
            if (du != inf) {
                b = rowptr[u];
                e = rowptr[u + 1];
            }
        }
        const bool light = (e - b) <= 32;
        if (light)
            for (int32_t i = b; i < e; ++i) {
                const int32_t v = col[i];
                if (v != u && v != src && du + w[i] == dist[v]) atomicMin(&pre[v], static_cast<int32_t>(u));
            }
        uint64_t heavy = __ballot(!light);
        while (heavy) {
            const int sl = __ffsll(static_cast<unsigned long long>(heavy)) - 1;
            heavy &= heavy - 1;
            const int32_t hb = __shfl(b, sl), he = __shfl(e, sl);
            const double hd = __shfl(du, sl);
            const int32_t hu = static_cast<int32_t>(g * 64 + sl);
            for (int32_t i = hb + lane; i < he; i += 64) {
                const int32_t v = col[i];
                if (v != hu && v != src && hd + w[i] == dist[v]) atomicMin(&pre[v], hu);
            }
        }
    }
}

__global__ void k_pre_finish(int32_t* pre, int32_t n) {
    for (int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < n;
         v += static_cast<int64_t>(gridDim.x) * blockDim.x)
        if (pre[v] == kIntInf) pre[v] = -1;
}

}  // namespace

// Base bucket width: ~32 average weights per average out-degree (GPU near-far
// favours wide buckets: a bucket should hold enough work to fill the chip).
double sssp_base_delta(const ghet_graph* g) {
    const double avg_deg = g->n > 0 ? static_cast<double>(g->out.m) / g->n : 1.0;
    return std::max(1e-12, 32.0 * 0.5 / std::max(avg_deg, 1.0));
}

int run_sssp(ghet_graph* g, int32_t src, double inf, double* dist_out, int32_t* pre_out, const ghet_opts* o,
             ghet_run_stats* st) {
    const int32_t n = g->n;
    if (src < 0 || src >= n) return set_error(GHET_ERR_INVALID, "source %d out of range [0,%d)", src, n);
    if (!g->out.present() || !g->out.w) return set_error(GHET_ERR_INVALID, "SSSP needs the weighted out-CSR");
    if (!(inf > 0) || std::isnan(inf)) return set_error(GHET_ERR_INVALID, "inf must be > 0");
    hipStream_t s = pick_stream(g, o);
    const bool have_opts = o && o->size >= sizeof(ghet_opts);
    const int schedule = have_opts ? o->schedule : GHET_SCHED_HEURISTIC;
    const double base_delta = (have_opts && o->sssp_delta > 0) ? o->sssp_delta : sssp_base_delta(g);
    ghet_level_trace* trace = have_opts ? o->trace : nullptr;
    const int trace_cap = trace ? o->trace_cap : 0;
    int trace_len = 0;

    double* dist;
    int32_t *qa, *qb, *scan, *aux, *stamp_near, *stamp_far, *far[2];
    SsspCtr* ctr;
    void* p;
    GHET_CHECK(g->ws.get(WS_DIST, 8 * size_t(n), &p)); dist = static_cast<double*>(p);
    GHET_CHECK(g->ws.get(WS_QUEUE_A, 4 * size_t(n), &p)); qa = static_cast<int32_t*>(p);
    GHET_CHECK(g->ws.get(WS_QUEUE_B, 4 * size_t(n), &p)); qb = static_cast<int32_t*>(p);
    GHET_CHECK(g->ws.get(WS_FAR_A, 4 * size_t(n), &p)); far[0] = static_cast<int32_t*>(p);
    GHET_CHECK(g->ws.get(WS_FAR_B, 4 * size_t(n), &p)); far[1] = static_cast<int32_t*>(p);
    GHET_CHECK(g->ws.get(WS_STAMP_NEAR, 4 * size_t(n), &p)); stamp_near = static_cast<int32_t*>(p);
    GHET_CHECK(g->ws.get(WS_STAMP_FAR, 4 * size_t(n), &p)); stamp_far = static_cast<int32_t*>(p);
    GHET_CHECK(g->ws.get(WS_SCAN, 4 * (size_t(n) + 1), &p)); scan = static_cast<int32_t*>(p);
    GHET_CHECK(g->ws.get(WS_SCAN_AUX, 4 * size_t(scan_aux_ints(n)), &p)); aux = static_cast<int32_t*>(p);
    GHET_CHECK(g->ws.get(WS_COUNTERS, 256, &p)); ctr = static_cast<SsspCtr*>(p);

    SchedState sched;
    sched_begin(&sched, schedule, GHET_NET_SSSP, n, g->out.m, g->out.m);

    GHET_HIP(hipEventRecord(g->ev_begin, s));
    GHET_CHECK(fill<double>(dist, inf, n, s));
    GHET_CHECK(fill<int32_t>(stamp_near, -1, n, s));
    GHET_CHECK(fill<int32_t>(stamp_far, -1, n, s));
    GHET_CHECK(fill<int32_t>(reinterpret_cast<int32_t*>(ctr), 0, 8, s));
    hipLaunchKernelGGL(k_sssp_seed, dim3(1), dim3(64), 0, s, dist, qa, stamp_near, src);

    int32_t F = 1;
    int64_t E = g->h_out_rowptr[src + 1] - g->h_out_rowptr[src];
    int32_t round = 0, far_gen = 0, phases = 0;
    int sel = 0;
    int32_t far_count = 0;
    double threshold = 0.0;
    double delta = base_delta;
    // first phase
    {
        LevelInfo li;
        li.n = n; li.m = g->out.m; li.m_in = g->out.m; li.level = 0; li.frontier = F; li.frontier_edges = E;
        li.unvisited = n - 1; li.delta = base_delta; li.threshold = 0; li.far_size = 0;
        int cls = sched_pick_sssp(&sched, li);
        if (trace && trace_len < trace_cap) {
            sched_features(li, trace[trace_len].features);
            trace[trace_len].mode = cls;
        }
        delta = base_delta * std::ldexp(1.0, cls - 1);
        threshold = delta;
    }
    int64_t rounds = 0;
    for (;;) {
        // relaxation rounds of the current bucket
        if (trace && trace_len < trace_cap) GHET_HIP(hipEventRecord(g->ev_k0, s));
        while (F > 0) {
            hipLaunchKernelGGL(k_round_reset, dim3(1), dim3(64), 0, s, ctr);
            hipLaunchKernelGGL(k_frontier_deg, dim3(static_cast<unsigned>(ceil_div(F, 256))), dim3(256), 0, s, qa, F,
                               g->out.rowptr, scan);
            GHET_CHECK(exclusive_scan_i32(scan, scan, F, aux, s));
            const int32_t next_round = round + 1;
            if (E > 0) {
                unsigned grid = static_cast<unsigned>(std::min<int64_t>(ceil_div(E, kTile), 65535));
                hipLaunchKernelGGL(k_relax, dim3(grid), dim3(kBT), 0, s, qa, F, scan, static_cast<int32_t>(E),
                                   g->out.rowptr, g->out.col, g->out.w, dist, threshold, stamp_near, next_round,
                                   stamp_far, far_gen, qb, far[sel], sel, ctr);
            }
            GHET_HIP(hipGetLastError());
            GHET_HIP(hipMemcpyAsync(g->pinned, ctr, sizeof(SsspCtr), hipMemcpyDeviceToHost, s));
            GHET_HIP(hipStreamSynchronize(s));
            const SsspCtr* hc = reinterpret_cast<const SsspCtr*>(g->pinned);
            F = hc->near_count;
            E = static_cast<int64_t>(hc->near_edges);
            far_count = hc->far_count[sel];
            std::swap(qa, qb);
            round = next_round;
            ++rounds;
        }
        if (trace && trace_len < trace_cap) {
            GHET_HIP(hipEventRecord(g->ev_k1, s));
            GHET_HIP(hipEventSynchronize(g->ev_k1));
            float ms = 0;
            GHET_HIP(hipEventElapsedTime(&ms, g->ev_k0, g->ev_k1));
            trace[trace_len].ms = ms;
            ++trace_len;
        }
        if (far_count == 0) break;
        ++phases;
        // advance the threshold and split the far list
        double old_thr = threshold;
        {
            LevelInfo li;
            li.n = n; li.m = g->out.m; li.m_in = g->out.m; li.level = phases; li.frontier = 0;
            li.frontier_edges = 0; li.unvisited = 0; li.delta = base_delta; li.threshold = threshold;
            li.far_size = far_count; li.prev_mode = 0;
            int cls = sched_pick_sssp(&sched, li);
            if (trace && trace_len < trace_cap) {
                sched_features(li, trace[trace_len].features);
                trace[trace_len].mode = cls;
            }
            delta = base_delta * std::ldexp(1.0, cls - 1);
        }
        double new_thr = std::max(threshold + delta, std::nextafter(threshold, HUGE_VAL));
        for (;;) {
            const int other = 1 - sel;
            hipLaunchKernelGGL(k_split_reset, dim3(1), dim3(64), 0, s, ctr, other);
            const int32_t next_round = round + 1;
            unsigned grid = static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(ceil_div(far_count, 256), 8192)));
            hipLaunchKernelGGL(k_far_split, dim3(grid), dim3(256), 0, s, far[sel], far_count, dist, old_thr, new_thr,
                               stamp_near, next_round, stamp_far, far_gen + 1, qa, far[other], other, g->out.rowptr,
                               ctr);
            GHET_HIP(hipGetLastError());
            GHET_HIP(hipMemcpyAsync(g->pinned, ctr, sizeof(SsspCtr), hipMemcpyDeviceToHost, s));
            GHET_HIP(hipStreamSynchronize(s));
            const SsspCtr* hc = reinterpret_cast<const SsspCtr*>(g->pinned);
            F = hc->near_count;
            E = static_cast<int64_t>(hc->near_edges);
            far_count = hc->far_count[other];
            double far_min;
            std::memcpy(&far_min, &hc->far_min_bits, sizeof(double));
            sel = other;
            ++far_gen;
            round = next_round;
            if (F > 0 || far_count == 0) break;
            // nothing in the next bucket: jump to the smallest far distance
            // (strictly above it, so the split always makes progress)
            old_thr = new_thr;
            new_thr = std::max(far_min + delta, std::nextafter(far_min, HUGE_VAL));
        }
        threshold = new_thr;
        if (F == 0 && far_count == 0) break;
    }
    if (pre_out) {
        GHET_CHECK(g->ws.get(WS_PRE, 4 * size_t(n), &p));
        int32_t* pre = static_cast<int32_t*>(p);
        GHET_CHECK(fill<int32_t>(pre, kIntInf, n, s));
        const unsigned gw = static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 4096)));
        hipLaunchKernelGGL(k_sssp_parent, dim3(gw), dim3(256), 0, s, n, g->out.rowptr, g->out.col, g->out.w, dist, inf,
                           src, pre);
        hipLaunchKernelGGL(k_pre_finish, dim3(static_cast<unsigned>(std::min<int64_t>(ceil_div(n, 256), 8192))),
                           dim3(256), 0, s, pre, n);
    }
    GHET_HIP(hipEventRecord(g->ev_end, s));
    GHET_HIP(hipGetLastError());
    if (dist_out) GHET_HIP(hipMemcpyAsync(dist_out, dist, 8 * size_t(n), hipMemcpyDeviceToHost, s));
    if (pre_out) GHET_HIP(hipMemcpyAsync(pre_out, g->ws.ptr[WS_PRE], 4 * size_t(n), hipMemcpyDeviceToHost, s));
    if (st) {
        GHET_CHECK(reached_stats_f64(dist, inf, g->out.rowptr, n, reinterpret_cast<int64_t*>(scan), s));
        GHET_HIP(hipMemcpyAsync(g->pinned, scan, 16, hipMemcpyDeviceToHost, s));
    }
    GHET_HIP(hipStreamSynchronize(s));
    if (st) {
        float ms = 0;
        GHET_HIP(hipEventElapsedTime(&ms, g->ev_begin, g->ev_end));
        st->device_ms = ms;
        st->levels = static_cast<int32_t>(rounds);
        const int64_t* rc = reinterpret_cast<const int64_t*>(g->pinned);
        st->vertices_reached = rc[0];
        st->edges_traversed = rc[1];
        st->trace_len = trace_len;
        st->extra[0] = base_delta;
        st->extra[1] = phases;
        st->status_flags = sched.fallback ? 1 : 0;
    }
    return sched.fallback ? GHET_WARN_FALLBACK : GHET_OK;
}

}  // namespace ghet
