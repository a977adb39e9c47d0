// Batched SSSP: K sources per pass advanced together, one launch per round
// for the whole batch (near-far buckets with a threshold shared by the batch).
//
// No reference counterpart: the reference runs one csr_sssp per source
// (GPU_SSSP_csr.cuh:253).  Every source converges to the same fixpoint as its
// own run (the minimum over walks of the rounded path sum), so results are
// bit-identical to per-source runs whatever the interleaving.
//
// Work items are (source, vertex) pairs packed as (s << 32) | v; distances
// and stamps are vertex-major [n][K]: the K sources' values of one vertex
// share cache lines, and R-MAT frontiers of different sources meet at the
// same hubs in the same rounds, so a pass of K sources touches far fewer
// lines than K separate searches.  The caller's [k][n] rows are produced by
// one tiled transpose at the end of a pass.  A round expands the concatenated near
// queues of all K sources edge-balanced (int64 degree scan, 1024-edge tiles),
// so even when each single search has a thin frontier the batch fills the chip.
#include <cmath>
#include <cstring>

#include "graph_internal.hpp"

namespace ghet {
double sssp_base_delta(const ghet_graph* g);

namespace {

struct MssCtr {
    unsigned long long near_count;
    unsigned long long far_count[2];
    unsigned long long pad[5];
    CtrSlot slots[kSlots];  // v[0] out-edges of next near items, v[1] bits of the smallest kept far distance
};
constexpr int kCtrWords = sizeof(MssCtr) / 8;
constexpr unsigned long long kInfBits = 0x7FF0000000000000ULL;

constexpr int kBT = 256;
constexpr int kItems = 4;
constexpr int kTile = kBT * kItems;
constexpr int kOut = 2 * kTile;
constexpr int kMaxBlocks = 2048;
constexpr int kMaxK = 256;     // sources per pass (vertex-major stride)
constexpr int kDefaultK = 16;

__device__ __forceinline__ int64_t upper_bound_i64(const int64_t* a, int64_t len, int64_t key) {
    int64_t lo = 0, hi = len;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (a[mid] <= key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ int32_t item_v(uint64_t it) { return static_cast<int32_t>(it & 0xFFFFFFFFu); }
__device__ __forceinline__ int64_t item_s(uint64_t it) { return static_cast<int64_t>(it >> 32); }

__global__ void k_mss_seed(const int32_t* __restrict__ srcs, int32_t k, int32_t n, double* dist, uint64_t* queue,
                           int32_t* stamp_near) {
    const int i = threadIdx.x;
    if (i < k) {
        const int64_t idx = static_cast<int64_t>(srcs[i]) * k + i;
        dist[idx] = 0.0;
        stamp_near[idx] = 0;
        queue[i] = (static_cast<uint64_t>(i) << 32) | static_cast<uint32_t>(srcs[i]);
    }
}

__global__ void k_mss_reset(MssCtr* c, int other) {
    unsigned long long* w = reinterpret_cast<unsigned long long*>(c->slots);
    for (int i = threadIdx.x; i < kSlots * 8; i += blockDim.x) w[i] = (i % 8 == 1) ? kInfBits : 0ULL;
    if (threadIdx.x == 0) {
        c->near_count = 0;
        if (other >= 0) c->far_count[other] = 0;
    }
}

__global__ void k_item_deg(const uint64_t* __restrict__ q, int64_t F, const int32_t* __restrict__ rowptr,
                           int64_t* __restrict__ deg) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < F;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int32_t v = item_v(q[i]);
        deg[i] = rowptr[v + 1] - rowptr[v];
    }
}

__global__ void __launch_bounds__(kBT) k_mss_relax(const uint64_t* __restrict__ q, int64_t F,
                                                   const int64_t* __restrict__ scan, int64_t E,
                                                   const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                                   const double* __restrict__ wts, int32_t ks, double* dist,
                                                   double threshold, int32_t* stamp_near, int32_t next_round,
                                                   int32_t* stamp_far, int32_t far_gen, uint64_t* __restrict__ next_q,
                                                   uint64_t* far_list, int far_sel, MssCtr* ctr) {
    __shared__ int64_t s_scan[kTile];
    __shared__ int32_t s_start[kTile];
    __shared__ uint64_t s_item[kTile];
    __shared__ uint64_t s_near[kOut], s_far[kOut];
    __shared__ int64_t s_i0, s_span;
    __shared__ int s_nc, s_fc;
    __shared__ unsigned long long s_base;
    __shared__ unsigned long long s_red[kBT / 64];
    const int tid = threadIdx.x;
    if (tid == 0) {
        s_nc = 0;
        s_fc = 0;
    }
    __syncthreads();  // counters zeroed before any wave appends
    unsigned long long od = 0;
    auto on_near = [&](uint64_t it) {
        const int32_t v = item_v(it);
        od += rowptr[v + 1] - rowptr[v];
    };
    auto nop = [](uint64_t) {};
    for (int64_t tile = blockIdx.x; tile * kTile < E; tile += gridDim.x) {
        const int64_t t0 = tile * kTile, t1 = std::min<int64_t>(t0 + kTile, E);
        if (tid == 0) {
            const int64_t i0 = upper_bound_i64(scan, F + 1, t0) - 1;
            const int64_t i1 = upper_bound_i64(scan, F + 1, t1 - 1) - 1;
            s_i0 = i0;
            s_span = i1 - i0 + 1;
        }
        __syncthreads();
        const int64_t i0 = s_i0, span = s_span;
        const bool use_lds = span <= kTile;
        if (use_lds) {
            for (int j = tid; j < span; j += kBT) {
                const uint64_t it = q[i0 + j];
                s_scan[j] = scan[i0 + j];
                s_start[j] = rowptr[item_v(it)];
                s_item[j] = it;
            }
        }
        __syncthreads();
        for (int k = 0; k < kItems; ++k) {
            const int64_t e = t0 + k * kBT + tid;
            bool to_near = false, to_far = false;
            uint64_t out = 0;
            if (e < t1) {
                int64_t pos;
                uint64_t it;
                if (use_lds) {
                    const int64_t j = upper_bound_i64(s_scan, span, e) - 1;
                    it = s_item[j];
                    pos = s_start[j] + (e - s_scan[j]);
                } else {
                    const int64_t j = upper_bound_i64(scan + i0, span, e) - 1;
                    it = q[i0 + j];
                    pos = rowptr[item_v(it)] + (e - scan[i0 + j]);
                }
                const int64_t sb = item_s(it);
                const int32_t v = col[pos];
                const int64_t pv = static_cast<int64_t>(v) * ks + sb;
                const double nd = dist[static_cast<int64_t>(item_v(it)) * ks + sb] + wts[pos];
                if (nd < dist[pv]) {
                    const double old = atomic_min_nonneg(&dist[pv], nd);
                    if (nd < old) {
                        if (nd < threshold) to_near = atomicExch(&stamp_near[pv], next_round) != next_round;
                        else to_far = atomicExch(&stamp_far[pv], far_gen) != far_gen;
                    }
                }
                out = (it & 0xFFFFFFFF00000000ULL) | static_cast<uint32_t>(v);
            }
            lq_push(s_near, &s_nc, to_near, out);
            lq_push(s_far, &s_fc, to_far, out);
        }
        lq_flush<kBT>(s_near, &s_nc, &s_base, &ctr->near_count, next_q, kOut - kTile, on_near);
        lq_flush<kBT>(s_far, &s_fc, &s_base, &ctr->far_count[far_sel], far_list, kOut - kTile, nop);
    }
    lq_flush<kBT>(s_near, &s_nc, &s_base, &ctr->near_count, next_q, 0, on_near);
    lq_flush<kBT>(s_far, &s_fc, &s_base, &ctr->far_count[far_sel], far_list, 0, nop);
    od = block_sum_u64<kBT>(od, s_red);
    if (tid == 0) slot_add(ctr->slots, 0, od);
}

__global__ void __launch_bounds__(kBT) k_mss_split(const uint64_t* __restrict__ far_in, int64_t count,
                                                   const double* __restrict__ dist, int32_t ks, double old_thr,
                                                   double new_thr, int32_t* stamp_near, int32_t next_round,
                                                   int32_t* stamp_far, int32_t new_gen, uint64_t* __restrict__ near_q,
                                                   uint64_t* far_out, int other, const int32_t* __restrict__ rowptr,
                                                   MssCtr* ctr) {
    __shared__ uint64_t s_near[kOut], s_far[kOut];
    __shared__ int s_nc, s_fc;
    __shared__ unsigned long long s_base;
    __shared__ unsigned long long s_red[kBT / 64];
    if (threadIdx.x == 0) {
        s_nc = 0;
        s_fc = 0;
    }
    __syncthreads();  // counters zeroed before any wave appends
    unsigned long long od = 0, dmin = kInfBits;
    auto on_near = [&](uint64_t it) {
        const int32_t v = item_v(it);
        od += rowptr[v + 1] - rowptr[v];
    };
    auto nop = [](uint64_t) {};
    for (int64_t base = static_cast<int64_t>(blockIdx.x) * kBT; base < count;
         base += static_cast<int64_t>(gridDim.x) * kBT) {
        const int64_t i = base + threadIdx.x;
        bool to_near = false, keep = false;
        uint64_t it = 0;
        if (i < count) {
            it = far_in[i];
            const int64_t idx = static_cast<int64_t>(item_v(it)) * ks + item_s(it);
            const double d = dist[idx];
            if (d >= old_thr) {
                if (d < new_thr) {
                    stamp_near[idx] = next_round;
                    to_near = true;
                } else {
                    stamp_far[idx] = new_gen;
                    keep = true;
                    const unsigned long long bits = static_cast<unsigned long long>(__double_as_longlong(d));
                    dmin = bits < dmin ? bits : dmin;
                }
            }
        }
        lq_push(s_near, &s_nc, to_near, it);
        lq_push(s_far, &s_fc, keep, it);
        lq_flush<kBT>(s_near, &s_nc, &s_base, &ctr->near_count, near_q, kOut - kBT, on_near);
        lq_flush<kBT>(s_far, &s_fc, &s_base, &ctr->far_count[other], far_out, kOut - kBT, nop);
    }
    lq_flush<kBT>(s_near, &s_nc, &s_base, &ctr->near_count, near_q, 0, on_near);
    lq_flush<kBT>(s_far, &s_fc, &s_base, &ctr->far_count[other], far_out, 0, nop);
    od = block_sum_u64<kBT>(od, s_red);
    dmin = wave_min(dmin);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = dmin;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 0; w < kBT / 64; ++w) dmin = s_red[w] < dmin ? s_red[w] : dmin;
        slot_add(ctr->slots, 0, od);
        if (dmin != kInfBits) atomicMin(&ctr->slots[blockIdx.x & (kSlots - 1)].v[1], dmin);
    }
}

// [n][ks] -> [ks][n] through a 64x64 LDS tile (coalesced on both sides).
__global__ void __launch_bounds__(256) k_mss_transpose(const double* __restrict__ in, int32_t n, int32_t ks,
                                                       double* __restrict__ out) {
    __shared__ double t[64][65];
    const int64_t v0 = static_cast<int64_t>(blockIdx.x) * 64, s0 = static_cast<int64_t>(blockIdx.y) * 64;
    for (int r = threadIdx.y; r < 64; r += 4) {
        const int64_t v = v0 + r, sidx = s0 + threadIdx.x;
        if (v < n && sidx < ks) t[r][threadIdx.x] = in[v * ks + sidx];
    }
    __syncthreads();
    for (int r = threadIdx.y; r < 64; r += 4) {
        const int64_t sidx = s0 + r, v = v0 + threadIdx.x;
        if (v < n && sidx < ks) out[sidx * n + v] = t[threadIdx.x][r];
    }
}

// Reached-vertex statistics over the batch: edges of (source, vertex) pairs
// with a finite distance.
__global__ void __launch_bounds__(256) k_mss_reached(const double* __restrict__ dist, int64_t total, int32_t ks,
                                                     double inf, const int32_t* __restrict__ rowptr, CtrSlot* slots) {
    __shared__ unsigned long long s_red[4];
    unsigned long long e = 0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        if (dist[i] != inf) {
            const int32_t v = static_cast<int32_t>(i / ks);
            e += rowptr[v + 1] - rowptr[v];
        }
    e = block_sum_u64<256>(e, s_red);
    if (threadIdx.x == 0) slot_add(slots, 2, e);
}

}  // namespace

int run_sssp_multi(ghet_graph* g, int32_t k, const int32_t* srcs, double inf, double* dist_out, const ghet_opts* o,
                   ghet_run_stats* st) {
    const int32_t n = g->n;
    if (k < 0 || (k > 0 && !srcs)) return set_error(GHET_ERR_INVALID, "bad source list");
    for (int32_t i = 0; i < k; ++i)
        if (srcs[i] < 0 || srcs[i] >= n) return set_error(GHET_ERR_INVALID, "source %d out of range [0,%d)", srcs[i], n);
    if (!g->out.present() || !g->out.w) return set_error(GHET_ERR_INVALID, "SSSP needs the weighted out-CSR");
    if (!(inf > 0) || std::isnan(inf)) return set_error(GHET_ERR_INVALID, "inf must be > 0");
    if (k == 0 || n == 0) return GHET_OK;
    hipStream_t s = pick_stream(g, o);
    const bool have_opts = o && o->size >= sizeof(ghet_opts);
    const double delta = (have_opts && o->sssp_delta > 0) ? o->sssp_delta : sssp_base_delta(g);
    // Sources per pass: 16 by default (fastest measured on R-MAT s22: 16 ->
    // 1692 ms, 64 -> 1944 ms, 256 -> 2140 ms for 256 sources; wider passes add
    // same-line atomic contention and rounds), at most kMaxK, bounded by a
    // 64 GB working set (dist, two stamps, four queues ~ 52 B per pair).
    const int64_t budget = int64_t(64) << 30;
    const int64_t kcap = (have_opts && o->reserved[3] > 0) ? o->reserved[3] : kDefaultK;
    const int32_t K = static_cast<int32_t>(
        std::max<int64_t>(1, std::min<int64_t>({kcap, int64_t(kMaxK), int64_t(k), budget / (52 * int64_t(n))})));
    const int64_t KN = int64_t(K) * n;
    void* p;
    double* dist;
    int32_t *stamp_near, *stamp_far, *d_src;
    uint64_t *qa, *qb, *far[2];
    int64_t *scan, *aux;
    MssCtr* ctr;
    GHET_CHECK(g->ws.get(WS_MULTI_A, 8 * size_t(KN), &p)); dist = static_cast<double*>(p);
    GHET_CHECK(g->ws.get(WS_MULTI_B, 8 * size_t(KN), &p));
    stamp_near = static_cast<int32_t*>(p);
    stamp_far = stamp_near + KN;
    GHET_CHECK(g->ws.get(WS_MULTI_C, 32 * size_t(KN), &p));
    qa = static_cast<uint64_t*>(p);
    qb = qa + KN;
    far[0] = qb + KN;
    far[1] = far[0] + KN;
    GHET_CHECK(g->ws.get(WS_MULTI_D, 8 * (size_t(KN) + 1), &p)); scan = static_cast<int64_t*>(p);
    GHET_CHECK(g->ws.get(WS_SCAN_AUX, 8 * size_t(scan_aux_ints(KN)), &p)); aux = static_cast<int64_t*>(p);
    GHET_CHECK(g->ws.get(WS_COUNTERS, sizeof(MssCtr), &p)); ctr = static_cast<MssCtr*>(p);
    GHET_CHECK(g->ws.get(WS_MISC, 4 * size_t(K) + 256, &p)); d_src = static_cast<int32_t*>(p);
    const MssCtr* hc = reinterpret_cast<const MssCtr*>(g->pinned);

    int64_t traversed = 0, rounds = 0;
    GHET_HIP(hipEventRecord(g->ev_begin, s));
    for (int32_t p0 = 0; p0 < k; p0 += K) {
        const int32_t kp = std::min(K, k - p0);
        const int64_t kn = int64_t(kp) * n;
        GHET_CHECK(fill<double>(dist, inf, kn, s));
        GHET_CHECK(fill<int32_t>(stamp_near, -1, kn, s));
        GHET_CHECK(fill<int32_t>(stamp_far, -1, kn, s));
        GHET_CHECK(fill<uint64_t>(reinterpret_cast<uint64_t*>(ctr), 0, kCtrWords, s));
        GHET_HIP(hipMemcpyAsync(d_src, srcs + p0, 4 * size_t(kp), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_mss_seed, dim3(1), dim3(kMaxK), 0, s, d_src, kp, n, dist, qa, stamp_near);
        int64_t F = kp, E = 0;
        for (int32_t i = 0; i < kp; ++i) E += g->h_out_rowptr[srcs[p0 + i] + 1] - g->h_out_rowptr[srcs[p0 + i]];
        int32_t round = 0, far_gen = 0;
        int sel = 0;
        int64_t far_count = 0;
        double threshold = delta;
        for (;;) {
            while (F > 0) {
                hipLaunchKernelGGL(k_mss_reset, dim3(1), dim3(512), 0, s, ctr, -1);
                hipLaunchKernelGGL(k_item_deg, dim3(static_cast<unsigned>(std::min<int64_t>(ceil_div(F, 256), 65535))),
                                   dim3(256), 0, s, qa, F, g->out.rowptr, scan);
                GHET_CHECK(exclusive_scan_i64(scan, scan, F, aux, s));
                const int32_t next_round = round + 1;
                if (E > 0) {
                    const unsigned grid = static_cast<unsigned>(std::min<int64_t>(ceil_div(E, kTile), kMaxBlocks));
                    hipLaunchKernelGGL(k_mss_relax, dim3(grid), dim3(kBT), 0, s, qa, F, scan, E, g->out.rowptr,
                                       g->out.col, g->out.w, kp, dist, threshold, stamp_near, next_round, stamp_far,
                                       far_gen, qb, far[sel], sel, ctr);
                }
                GHET_HIP(hipGetLastError());
                GHET_HIP(hipMemcpyAsync(g->pinned, ctr, sizeof(MssCtr), hipMemcpyDeviceToHost, s));
                GHET_HIP(hipStreamSynchronize(s));
                F = static_cast<int64_t>(hc->near_count);
                E = static_cast<int64_t>(slot_sum(hc->slots, 0));
                far_count = static_cast<int64_t>(hc->far_count[sel]);
                std::swap(qa, qb);
                round = next_round;
                ++rounds;
            }
            if (far_count == 0) break;
            double old_thr = threshold;
            double new_thr = std::max(threshold + delta, std::nextafter(threshold, HUGE_VAL));
            for (;;) {
                const int other = 1 - sel;
                hipLaunchKernelGGL(k_mss_reset, dim3(1), dim3(512), 0, s, ctr, other);
                const int32_t next_round = round + 1;
                const unsigned grid = static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(ceil_div(far_count, kBT), 2048)));
                hipLaunchKernelGGL(k_mss_split, dim3(grid), dim3(kBT), 0, s, far[sel], far_count, dist, kp, old_thr,
                                   new_thr, stamp_near, next_round, stamp_far, far_gen + 1, qa, far[other], other,
                                   g->out.rowptr, ctr);
                GHET_HIP(hipGetLastError());
                GHET_HIP(hipMemcpyAsync(g->pinned, ctr, sizeof(MssCtr), hipMemcpyDeviceToHost, s));
                GHET_HIP(hipStreamSynchronize(s));
                F = static_cast<int64_t>(hc->near_count);
                E = static_cast<int64_t>(slot_sum(hc->slots, 0));
                far_count = static_cast<int64_t>(hc->far_count[other]);
                unsigned long long mb = kInfBits;
                for (int i = 0; i < kSlots; ++i) mb = std::min(mb, hc->slots[i].v[1]);
                double far_min;
                std::memcpy(&far_min, &mb, sizeof(double));
                sel = other;
                ++far_gen;
                round = next_round;
                if (F > 0 || far_count == 0) break;
                old_thr = new_thr;
                new_thr = std::max(far_min + delta, std::nextafter(far_min, HUGE_VAL));
            }
            threshold = new_thr;
            if (F == 0 && far_count == 0) break;
        }
        if (st) {
            hipLaunchKernelGGL(k_mss_reset, dim3(1), dim3(512), 0, s, ctr, -1);
            hipLaunchKernelGGL(k_mss_reached, dim3(2048), dim3(256), 0, s, dist, kn, kp, inf, g->out.rowptr, ctr->slots);
            GHET_HIP(hipMemcpyAsync(g->pinned, ctr, sizeof(MssCtr), hipMemcpyDeviceToHost, s));
        }
        if (dist_out) {
            // vertex-major [n][kp] -> the caller's source-major [kp][n], staged in the (now idle) queue
            double* tr = reinterpret_cast<double*>(qa);
            hipLaunchKernelGGL(k_mss_transpose, dim3(static_cast<unsigned>(ceil_div(n, 64)), static_cast<unsigned>(ceil_div(kp, 64))),
                               dim3(64, 4), 0, s, dist, n, kp, tr);
            GHET_HIP(hipGetLastError());
            GHET_HIP(hipMemcpyAsync(dist_out + int64_t(p0) * n, tr, 8 * size_t(kn), hipMemcpyDeviceToHost, s));
        }
        GHET_HIP(hipStreamSynchronize(s));
        if (st) traversed += static_cast<int64_t>(slot_sum(hc->slots, 2));
    }
    GHET_HIP(hipEventRecord(g->ev_end, s));
    GHET_HIP(hipStreamSynchronize(s));
    if (st) {
        float ms = 0;
        GHET_HIP(hipEventElapsedTime(&ms, g->ev_begin, g->ev_end));
        st->device_ms = ms;
        st->levels = static_cast<int32_t>(rounds);
        st->edges_traversed = traversed;
        st->extra[0] = delta;
        st->extra[1] = K;
    }
    return GHET_OK;
}

}  // namespace ghet
