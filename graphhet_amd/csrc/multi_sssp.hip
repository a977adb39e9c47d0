// Batched SSSP: up to 32 sources per pass advanced together, one launch per
// round for the whole batch (near-far buckets with a threshold shared by the
// batch, light/heavy edge split as in the single-source form).
//
// No reference counterpart: the reference runs one csr_sssp per source
// (GPU_SSSP_csr.cuh:253).  Every source converges to the same fixpoint as its
// own run (the minimum over walks of the left-to-right rounded path sum, as
// for a single source), so results are bit-identical to per-source runs
// whatever the interleaving.
//
// Vertex queues with source masks.  A round's frontier is a queue of DISTINCT
// vertices; vertex u carries a 32-bit mask of the sources whose distance at
// u improved below the threshold.  One scan over the queue balances the
// round's edges for all sources at once, and each edge (u -> v, w) is read
// once and relaxed for every source in u's mask.  Distances are vertex-major
// [n][K] (the K values of a vertex share cache lines).  Rows are ordered by
// weight (the single-source layout), so a round reads only the prefix of u's
// row that is light for the smallest masked distance, and relaxes per source
// only the edges with d_s[u] + w < thr; an improvement ORs the source's bit
// into v's next-round mask, whose 0 -> nonzero transition appends v to the
// next queue.  It also ORs the bit into v's bucket mask, whose 0 -> nonzero
// transition appends v to the bucket list.  When the near queue drains, one
// heavy pass over the bucket list relaxes, per bucket-mask source, the edges
// with d_s[u] + w >= thr (now final) and sends improvements to the far list
// through far masks.  The far list is then split at the next threshold per
// (vertex, source) bit; bits of entries already below the old threshold are
// stale and dropped.
//
// The caller's [k][n] rows are produced by one tiled transpose per pass.
#include <cmath>
#include <cstring>

#include "graph_internal.hpp"

namespace ghet {

int sssp_base_delta(ghet_graph* g, double* out, bool size_term);
int build_sssp_layout(ghet_graph* g);

namespace {

// v[0] out-edges of the next near queue, v[1] bits of the smallest kept far
// distance, v[2] out-edges appended to the bucket list, v[3] edges scanned.
struct MssCtr {
    int32_t near_count;
    int32_t far_count[2];
    int32_t list_count;
    int32_t pad[12];
    CtrSlot slots[kSlots];
};
constexpr int kCtrWords = sizeof(MssCtr) / 8;
constexpr unsigned long long kInfBits = 0x7FF0000000000000ULL;

constexpr int kBT = 256;
constexpr int kItems = 8;
constexpr int kTile = kBT * kItems;
// LDS capacity of each staged queue (a tile pushes at most kTile entries; a
// flush empties it above kOut - kTile): 1.25 tiles instead of 2 lets 3 relax
// blocks share a CU instead of 2 (256 sources on s22: 483 -> 456 ms).
constexpr int kOut = kTile + kTile / 4;
// a relax tile pushes at most kTile entries per queue, a split step kBT
static_assert(kOut >= kTile && kOut >= kBT, "staged queues smaller than one step's pushes");
constexpr int kMaxBlocks = 2048;
constexpr int kMaxK = 32;  // sources per pass (mask width)

__device__ __forceinline__ int upper_bound_i32(const int32_t* a, int len, int32_t key) {
    int lo = 0, hi = len;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (a[mid] <= key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Seeds: q[j] = distinct source vertex with its sources qmask[j]: distance 0,
// near and bucket masks, and the bucket list (= the first near queue).
__global__ void k_mss_seed(const int32_t* __restrict__ q, const uint32_t* __restrict__ qmask, int32_t F, int32_t ks,
                           double* dist, uint32_t* mask, uint32_t* bmask, int32_t* list, MssCtr* ctr) {
    for (int j = threadIdx.x; j < F; j += blockDim.x) {
        const int32_t v = q[j];
        const uint32_t m = qmask[j];
        mask[v] = m;
        bmask[v] = m;
        list[j] = v;
        for (uint32_t b = m; b; b &= b - 1) dist[static_cast<int64_t>(v) * ks + (__ffs(b) - 1)] = 0.0;
    }
    if (threadIdx.x == 0) ctr->list_count = F;
}

// Resets the near queue and the spread sums; also far list `other` (split)
// and the bucket list (split: a new bucket starts).
__global__ void k_mss_reset(MssCtr* c, int other) {
    unsigned long long* w = reinterpret_cast<unsigned long long*>(c->slots);
    for (int i = threadIdx.x; i < kSlots * 8; i += blockDim.x) w[i] = (i % 8 == 1) ? kInfBits : 0ULL;
    if (threadIdx.x == 0) {
        c->near_count = 0;
        if (other >= 0) {
            c->far_count[other] = 0;
            c->list_count = 0;
        }
    }
}

// Per queue entry u with source mask m: the row segment a pass reads.  Round:
// the prefix with min_s d_s[u] + w < thr (it holds every source's light
// edges); heavy: the rest from the prefix end of max_s d_s[u] (it holds every
// source's heavy edges).
template <bool kHeavy>
__global__ void k_mss_rows(const int32_t* __restrict__ q, int32_t F, const uint32_t* __restrict__ qmask,
                           const int32_t* __restrict__ rowptr, const double* __restrict__ w,
                           const double* __restrict__ dist, int32_t ks, double threshold, int32_t* __restrict__ cnt,
                           int32_t* __restrict__ start) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < F;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int32_t u = q[i];
        const uint32_t m = qmask[u];
        const double* du = dist + static_cast<int64_t>(u) * ks;
        double d = kHeavy ? 0.0 : HUGE_VAL;
        for (uint32_t b = m; b; b &= b - 1) {
            const double x = du[__ffs(b) - 1];
            d = kHeavy ? fmax(d, x) : fmin(d, x);
        }
        const int32_t rb = rowptr[u], re = rowptr[u + 1];
        int32_t lo = rb, hi = re;
        while (lo < hi) {
            const int32_t mid = (lo + hi) >> 1;
            if (d + w[mid] < threshold) lo = mid + 1;
            else hi = mid;
        }
        cnt[i] = kHeavy ? re - lo : lo - rb;
        start[i] = kHeavy ? lo : rb;
    }
}

// Small queues (at most kSmallMax entries): one 1024-thread block resets the
// pass's counters (as k_mss_reset with no far list), computes the row
// segments (as k_mss_rows) and scans the counts.
constexpr int kSmallBT = 1024;
constexpr int kSmallItems = 4;
constexpr int kSmallMax = kSmallBT * kSmallItems;

template <bool kHeavy>
__global__ void __launch_bounds__(kSmallBT) k_mss_rows_small(const int32_t* __restrict__ q, int32_t F,
                                                             const uint32_t* __restrict__ qmask,
                                                             const int32_t* __restrict__ rowptr,
                                                             const double* __restrict__ w,
                                                             const double* __restrict__ dist, int32_t ks,
                                                             double threshold, int32_t* __restrict__ scan,
                                                             int32_t* __restrict__ start, MssCtr* ctr) {
    __shared__ int32_t s_wave[kSmallBT / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    unsigned long long* words = reinterpret_cast<unsigned long long*>(ctr->slots);
    for (int i = tid; i < kSlots * 8; i += kSmallBT) words[i] = (i % 8 == 1) ? kInfBits : 0ULL;
    if (tid == 0) ctr->near_count = 0;
    int32_t c[kSmallItems], st[kSmallItems];
    int32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kSmallItems; ++k) {
        const int i = tid * kSmallItems + k;
        c[k] = 0;
        st[k] = 0;
        if (i < F) {
            const int32_t u = q[i];
            const uint32_t m = qmask[u];
            const double* du = dist + static_cast<int64_t>(u) * ks;
            double d = kHeavy ? 0.0 : HUGE_VAL;
            for (uint32_t b = m; b; b &= b - 1) {
                const double x = du[__ffs(b) - 1];
                d = kHeavy ? fmax(d, x) : fmin(d, x);
            }
            const int32_t rb = rowptr[u], re = rowptr[u + 1];
            int32_t lo = rb, hi = re;
            while (lo < hi) {
                const int32_t mid = (lo + hi) >> 1;
                if (d + w[mid] < threshold) lo = mid + 1;
                else hi = mid;
            }
            c[k] = kHeavy ? re - lo : lo - rb;
            st[k] = kHeavy ? lo : rb;
        }
        sum += c[k];
    }
    int32_t incl = sum;
    for (int off = 1; off < 64; off <<= 1) {
        const int32_t y = __shfl_up(incl, off);
        if (lane >= off) incl += y;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    int32_t run = incl - sum, total = 0;
    for (int v = 0; v < kSmallBT / 64; ++v) {
        if (v < wave) run += s_wave[v];
        total += s_wave[v];
    }
#pragma unroll
    for (int k = 0; k < kSmallItems; ++k) {
        const int i = tid * kSmallItems + k;
        if (i < F) {
            scan[i] = run;
            start[i] = st[k];
            run += c[k];
        }
    }
    if (tid == 0) scan[F] = total;
}

// The queue's masks are consumed: clear them for reuse two rounds later.
__global__ void k_clear_mask(const int32_t* __restrict__ q, int32_t F, uint32_t* __restrict__ mask) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < F;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        mask[q[i]] = 0u;
}

// One pass, persistent blocks over 2048-edge tiles; entry j of queue q covers
// edges [start[j], start[j] + scan[j+1] - scan[j]) of the weight-ordered rows
// and scan[F] is the pass's edge count.
// Round (kHeavy false): per source in mask_cur[u], relaxes the edges with
// d_s[u] + w < thr; improvements go to v's next-round mask (0 -> nonzero
// appends v to next_q) and its bucket mask (0 -> nonzero appends v to list).
// Heavy (kHeavy true): q is the drained bucket's list and mask_cur the
// bucket masks; relaxes the edges with d_s[u] + w >= thr, improvements go to
// v's far mask (0 -> nonzero appends v to the far list).
template <bool kHeavy>
__global__ void __launch_bounds__(kBT) k_mss_relax(const int32_t* __restrict__ q, int32_t F,
                                                   const int32_t* __restrict__ scan,
                                                   const int32_t* __restrict__ start,
                                                   const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                                   const double* __restrict__ wts, int32_t ks, double* dist,
                                                   double threshold, const uint32_t* __restrict__ mask_cur,
                                                   uint32_t* mask_next, uint32_t* bmask, uint32_t* fmask,
                                                   int32_t* __restrict__ next_q, int32_t* __restrict__ list,
                                                   int32_t* far_list, int far_sel, MssCtr* ctr) {
    __shared__ int32_t s_scan[kTile];
    __shared__ int32_t s_start[kTile];
    __shared__ int32_t s_src[kTile];
    __shared__ uint32_t s_mask[kTile];
    __shared__ int32_t s_qa[kOut], s_qb[kOut];  // round: near, list; heavy: far
    __shared__ int32_t s_i0, s_span, s_na, s_nb, s_base;
    __shared__ unsigned long long s_red[kBT / 64];
    const int tid = threadIdx.x;
    if (tid == 0) {
        s_na = 0;
        s_nb = 0;
    }
    __syncthreads();  // counters zeroed before any wave appends
    unsigned long long od = 0, ld = 0;
    auto on_near = [&](int32_t v) { od += rowptr[v + 1] - rowptr[v]; };
    auto on_list = [&](int32_t v) { ld += rowptr[v + 1] - rowptr[v]; };
    auto nop = [](int32_t) {};
    const int32_t E = scan[F];
    if (blockIdx.x == 0 && tid == 0) slot_add(ctr->slots, 3, static_cast<unsigned long long>(E));
    for (int64_t tile = blockIdx.x; tile * kTile < E; tile += gridDim.x) {
        const int32_t t0 = static_cast<int32_t>(tile * kTile);
        const int32_t t1 = static_cast<int32_t>(std::min<int64_t>(static_cast<int64_t>(t0) + kTile, E));
        if (tid == 0) {
            const int i0 = upper_bound_i32(scan, F + 1, t0) - 1;
            const int i1 = upper_bound_i32(scan, F + 1, t1 - 1) - 1;
            s_i0 = i0;
            s_span = i1 - i0 + 1;
        }
        __syncthreads();
        const int i0 = s_i0, span = s_span;
        const bool use_lds = span <= kTile;
        if (use_lds)
            for (int j = tid; j < span; j += kBT) {
                const int32_t u = q[i0 + j];
                s_scan[j] = scan[i0 + j];
                s_start[j] = start[i0 + j];
                s_src[j] = u;
                s_mask[j] = mask_cur[u];
            }
        __syncthreads();
        // Each stage runs over the thread's kItems edges together, so their
        // row, weight and target loads are in flight at the same time.
        int32_t u[kItems], pos[kItems], v[kItems];
        uint32_t msk[kItems];
        double w[kItems];
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            const int32_t e = t0 + k * kBT + tid;
            u[k] = 0;
            pos[k] = 0;  // E > 0, so edge 0 exists: masked lanes load it
            msk[k] = 0;
            if (e < t1) {
                if (use_lds) {
                    const int j = upper_bound_i32(s_scan, span, e) - 1;
                    pos[k] = s_start[j] + (e - s_scan[j]);
                    u[k] = s_src[j];
                    msk[k] = s_mask[j];
                } else {
                    const int j = upper_bound_i32(scan + i0, span, e) - 1;
                    u[k] = q[i0 + j];
                    pos[k] = start[i0 + j] + (e - scan[i0 + j]);
                    msk[k] = mask_cur[u[k]];
                }
            }
        }
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            v[k] = col[pos[k]];
            w[k] = wts[pos[k]];
        }
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            bool push_a = false, push_b = false;
            const double* du = dist + static_cast<int64_t>(u[k]) * ks;
            double* dv = dist + static_cast<int64_t>(v[k]) * ks;
            uint32_t bits = 0;
            for (uint32_t b = msk[k]; b; b &= b - 1) {
                const int s = __ffs(b) - 1;
                const double nd = du[s] + w[k];
                if ((kHeavy ? nd >= threshold : nd < threshold) && nd < dv[s]) {
                    const double old = atomic_min_nonneg(&dv[s], nd);
                    if (nd < old) bits |= 1u << s;
                }
            }
            if (bits) {
                if (kHeavy) {
                    push_a = atomicOr(&fmask[v[k]], bits) == 0u;
                } else {
                    push_a = atomicOr(&mask_next[v[k]], bits) == 0u;
                    push_b = atomicOr(&bmask[v[k]], bits) == 0u;
                }
            }
            lq_push(s_qa, &s_na, push_a, v[k]);
            if (!kHeavy) lq_push(s_qb, &s_nb, push_b, v[k]);
        }
        if (kHeavy) {
            lq_flush<kBT>(s_qa, &s_na, &s_base, &ctr->far_count[far_sel], far_list, kOut - kTile, nop);
        } else {
            lq_flush<kBT>(s_qa, &s_na, &s_base, &ctr->near_count, next_q, kOut - kTile, on_near);
            lq_flush<kBT>(s_qb, &s_nb, &s_base, &ctr->list_count, list, kOut - kTile, on_list);
        }
    }
    if (kHeavy) {
        lq_flush<kBT>(s_qa, &s_na, &s_base, &ctr->far_count[far_sel], far_list, 0, nop);
    } else {
        lq_flush<kBT>(s_qa, &s_na, &s_base, &ctr->near_count, next_q, 0, on_near);
        lq_flush<kBT>(s_qb, &s_nb, &s_base, &ctr->list_count, list, 0, on_list);
        od = block_sum_u64<kBT>(od, s_red);
        ld = block_sum_u64<kBT>(ld, s_red);
        if (tid == 0) {
            slot_add(ctr->slots, 0, od);
            slot_add(ctr->slots, 2, ld);
        }
    }
}

// Splits far list `far_in` at new_thr: per (vertex, source) bit, entries below
// old_thr are stale (settled in an earlier bucket) and dropped, entries below
// new_thr start the next bucket, the rest stay far.  A vertex with bits in
// the new bucket is appended to the bucket list, which is also the bucket's
// first near queue (its near and bucket masks get those bits; both masks are
// all-zero here and far-list entries are distinct).
__global__ void __launch_bounds__(kBT) k_mss_split(const int32_t* __restrict__ far_in, int32_t count,
                                                   const double* __restrict__ dist, int32_t ks, double old_thr,
                                                   double new_thr, uint32_t* fmask, uint32_t* mask_near,
                                                   uint32_t* bmask, int32_t* __restrict__ list, int32_t* far_out,
                                                   int other, const int32_t* __restrict__ rowptr, MssCtr* ctr) {
    __shared__ int32_t s_near[kOut], s_far[kOut];
    __shared__ int32_t s_nc, s_fc, s_base;
    __shared__ unsigned long long s_red[kBT / 64];
    if (threadIdx.x == 0) {
        s_nc = 0;
        s_fc = 0;
    }
    __syncthreads();  // counters zeroed before any wave appends
    unsigned long long od = 0, dmin = kInfBits;
    auto on_near = [&](int32_t v) { od += rowptr[v + 1] - rowptr[v]; };
    auto nop = [](int32_t) {};
    for (int64_t base = static_cast<int64_t>(blockIdx.x) * kBT; base < count;
         base += static_cast<int64_t>(gridDim.x) * kBT) {
        const int64_t i = base + threadIdx.x;
        bool to_near = false, keep = false;
        int32_t v = 0;
        if (i < count) {
            v = far_in[i];
            const uint32_t m = fmask[v];
            const double* dv = dist + static_cast<int64_t>(v) * ks;
            uint32_t nb = 0, kb = 0;
            for (uint32_t b = m; b; b &= b - 1) {
                const int s = __ffs(b) - 1;
                const double d = dv[s];
                if (d >= old_thr) {
                    if (d < new_thr) nb |= 1u << s;
                    else {
                        kb |= 1u << s;
                        const unsigned long long bits = static_cast<unsigned long long>(__double_as_longlong(d));
                        dmin = bits < dmin ? bits : dmin;
                    }
                }
            }
            fmask[v] = kb;
            keep = kb != 0u;
            if (nb) {
                mask_near[v] = nb;
                bmask[v] = nb;
                to_near = true;
            }
        }
        lq_push(s_near, &s_nc, to_near, v);
        lq_push(s_far, &s_fc, keep, v);
        lq_flush<kBT>(s_near, &s_nc, &s_base, &ctr->list_count, list, kOut - kBT, on_near);
        lq_flush<kBT>(s_far, &s_fc, &s_base, &ctr->far_count[other], far_out, kOut - kBT, nop);
    }
    lq_flush<kBT>(s_near, &s_nc, &s_base, &ctr->list_count, list, 0, on_near);
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

// Reached-vertex statistics over the batch: out-edges of (source, vertex)
// pairs with a finite distance.
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
    // A round carries up to 32 sources, so twice the single-source base
    // width without its small-graph term fills the chip (measured best at
    // scale 22, 256 sources).
    double delta = 0;
    if (have_opts && o->sssp_delta > 0) delta = o->sssp_delta;
    else {
        GHET_CHECK(sssp_base_delta(g, &delta, false));
        delta *= 2.0;
    }
    // Sources per pass: up to 32 (the mask width), or the caller's choice.
    const int64_t kcap = (have_opts && o->reserved[3] > 0) ? o->reserved[3] : kMaxK;
    const int32_t K = static_cast<int32_t>(std::max<int64_t>(1, std::min<int64_t>({kcap, int64_t(kMaxK), int64_t(k)})));
    const int64_t KN = int64_t(K) * n;
    void* p;
    double* dist;
    uint32_t *mask_a, *mask_b, *fmask, *bmask, *d_qmask;
    int32_t *qa, *qb, *far[2], *list, *qstart, *scan, *aux, *d_q;
    MssCtr* ctr;
    GHET_CHECK(build_sssp_layout(g));
    const SsspLayout& L = g->sl;
    GHET_CHECK(g->ws.get(WS_MULTI_A, 8 * size_t(KN), &p)); dist = static_cast<double*>(p);
    GHET_CHECK(g->ws.get(WS_MULTI_B, 16 * size_t(n), &p));
    mask_a = static_cast<uint32_t*>(p);
    mask_b = mask_a + n;
    fmask = mask_b + n;
    bmask = fmask + n;
    GHET_CHECK(g->ws.get(WS_MULTI_C, 24 * size_t(n), &p));
    qa = static_cast<int32_t*>(p);
    qb = qa + n;
    far[0] = qb + n;
    far[1] = far[0] + n;
    list = far[1] + n;
    qstart = list + n;
    GHET_CHECK(g->ws.get(WS_MULTI_D, 4 * (size_t(n) + 1), &p)); scan = static_cast<int32_t*>(p);
    GHET_CHECK(g->ws.get(WS_SCAN_AUX, 4 * size_t(scan_aux_ints(n)), &p)); aux = static_cast<int32_t*>(p);
    GHET_CHECK(g->ws.get(WS_COUNTERS, sizeof(MssCtr), &p)); ctr = static_cast<MssCtr*>(p);
    GHET_CHECK(g->ws.get(WS_MISC, 8 * size_t(kMaxK) + 256, &p));
    d_q = static_cast<int32_t*>(p);
    d_qmask = reinterpret_cast<uint32_t*>(d_q + kMaxK);
    double* tr = nullptr;
    if (dist_out) {
        GHET_CHECK(g->ws.get(WS_PR_OUT, 8 * size_t(KN), &p));
        tr = static_cast<double*>(p);
    }
    const MssCtr* hc = reinterpret_cast<const MssCtr*>(g->pinned);

    int64_t traversed = 0, rounds = 0, scanned = 0;
    GHET_HIP(hipEventRecord(g->ev_begin, s));
    for (int32_t p0 = 0; p0 < k; p0 += K) {
        const int32_t kp = std::min(K, k - p0);
        const int64_t kn = int64_t(kp) * n;
        GHET_CHECK(fill<double>(dist, inf, kn, s));
        GHET_CHECK(fill<uint32_t>(mask_a, 0u, 4 * int64_t(n), s));
        GHET_CHECK(fill<uint64_t>(reinterpret_cast<uint64_t*>(ctr), 0, kCtrWords, s));
        // the pass's distinct source vertices and their source masks
        std::vector<int32_t> hq;
        std::vector<uint32_t> hm;
        for (int32_t i = 0; i < kp; ++i) {
            const int32_t v = srcs[p0 + i];
            size_t j = 0;
            while (j < hq.size() && hq[j] != v) ++j;
            if (j == hq.size()) {
                hq.push_back(v);
                hm.push_back(0u);
            }
            hm[j] |= 1u << i;
        }
        int32_t F = static_cast<int32_t>(hq.size());
        int64_t E = 0;
        for (int32_t v : hq) E += g->h_out_rowptr[v + 1] - g->h_out_rowptr[v];
        int32_t list_len = F;
        int64_t list_edges = E;
        // pageable sources: hipMemcpyAsync stages them before returning
        GHET_HIP(hipMemcpyAsync(d_q, hq.data(), 4 * hq.size(), hipMemcpyHostToDevice, s));
        GHET_HIP(hipMemcpyAsync(d_qmask, hm.data(), 4 * hm.size(), hipMemcpyHostToDevice, s));
        uint32_t *mcur = mask_a, *mnext = mask_b;
        hipLaunchKernelGGL(k_mss_seed, dim3(1), dim3(64), 0, s, d_q, d_qmask, F, kp, dist, mcur, bmask, list, ctr);
        // a bucket's first near queue is its list; later rounds alternate qa/qb
        const int32_t* cur = list;
        int32_t* nxt = qa;
        int sel = 0;
        int64_t far_count = 0;
        double threshold = delta;
        // Edge-balanced pass over queue q (F entries, masks qm, rows holding
        // at most E_max edges: the grid bound).
        // The pass also resets the counters it fills.
        auto pass = [&](bool heavy, const int32_t* q, int32_t Fq, const uint32_t* qm, int64_t E_max) -> int {
            const unsigned vgrid = static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(ceil_div(Fq, 256), 65535)));
            if (Fq <= kSmallMax) {
                if (heavy)
                    hipLaunchKernelGGL(k_mss_rows_small<true>, dim3(1), dim3(kSmallBT), 0, s, q, Fq, qm, g->out.rowptr,
                                       L.w, dist, kp, threshold, scan, qstart, ctr);
                else
                    hipLaunchKernelGGL(k_mss_rows_small<false>, dim3(1), dim3(kSmallBT), 0, s, q, Fq, qm, g->out.rowptr,
                                       L.w, dist, kp, threshold, scan, qstart, ctr);
            } else {
                hipLaunchKernelGGL(k_mss_reset, dim3(1), dim3(512), 0, s, ctr, -1);
                if (heavy)
                    hipLaunchKernelGGL(k_mss_rows<true>, dim3(vgrid), dim3(256), 0, s, q, Fq, qm, g->out.rowptr, L.w,
                                       dist, kp, threshold, scan, qstart);
                else
                    hipLaunchKernelGGL(k_mss_rows<false>, dim3(vgrid), dim3(256), 0, s, q, Fq, qm, g->out.rowptr, L.w,
                                       dist, kp, threshold, scan, qstart);
                GHET_CHECK(exclusive_scan_i32(scan, scan, Fq, aux, s));
            }
            const unsigned grid = static_cast<unsigned>(std::min<int64_t>(ceil_div(E_max, kTile), kMaxBlocks));
            if (heavy)
                hipLaunchKernelGGL(k_mss_relax<true>, dim3(grid), dim3(kBT), 0, s, q, Fq, scan, qstart, g->out.rowptr,
                                   L.col, L.w, kp, dist, threshold, qm, mnext, bmask, fmask, nxt, list, far[sel], sel,
                                   ctr);
            else
                hipLaunchKernelGGL(k_mss_relax<false>, dim3(grid), dim3(kBT), 0, s, q, Fq, scan, qstart, g->out.rowptr,
                                   L.col, L.w, kp, dist, threshold, qm, mnext, bmask, fmask, nxt, list, far[sel], sel,
                                   ctr);
            // the pass's masks are consumed: clear them for reuse
            hipLaunchKernelGGL(k_clear_mask, dim3(vgrid), dim3(256), 0, s, q, Fq, const_cast<uint32_t*>(qm));
            GHET_HIP(hipGetLastError());
            GHET_HIP(hipMemcpyAsync(g->pinned, ctr, sizeof(MssCtr), hipMemcpyDeviceToHost, s));
            GHET_HIP(hipStreamSynchronize(s));
            scanned += static_cast<int64_t>(slot_sum(hc->slots, 3));
            return GHET_OK;
        };
        for (;;) {
            while (F > 0) {
                if (E > 0) {
                    GHET_CHECK(pass(false, cur, F, mcur, E));
                    F = hc->near_count;
                    E = static_cast<int64_t>(slot_sum(hc->slots, 0));
                    list_edges += static_cast<int64_t>(slot_sum(hc->slots, 2));
                    list_len = hc->list_count;
                } else {
                    hipLaunchKernelGGL(k_clear_mask, dim3(static_cast<unsigned>(ceil_div(F, 256))), dim3(256), 0, s,
                                       cur, F, mcur);
                    F = 0;
                }
                cur = nxt;
                nxt = (nxt == qa) ? qb : qa;
                std::swap(mcur, mnext);
                ++rounds;
            }
            // heavy pass over the drained bucket (also clears its bucket masks)
            if (list_len > 0) {
                if (list_edges > 0) {
                    GHET_CHECK(pass(true, list, list_len, bmask, list_edges));
                    far_count = hc->far_count[sel];
                } else {
                    hipLaunchKernelGGL(k_clear_mask, dim3(static_cast<unsigned>(ceil_div(list_len, 256))), dim3(256),
                                       0, s, list, list_len, bmask);
                }
            }
            if (far_count == 0) break;
            // split the far list at the next threshold into a new bucket list
            // (its near masks go to mcur, all-zero after the last round)
            double old_thr = threshold;
            double new_thr = std::max(threshold + delta, std::nextafter(threshold, HUGE_VAL));
            for (;;) {
                const int other = 1 - sel;
                hipLaunchKernelGGL(k_mss_reset, dim3(1), dim3(512), 0, s, ctr, other);
                const unsigned grid =
                    static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(ceil_div(far_count, kBT), 2048)));
                hipLaunchKernelGGL(k_mss_split, dim3(grid), dim3(kBT), 0, s, far[sel], static_cast<int32_t>(far_count),
                                   dist, kp, old_thr, new_thr, fmask, mcur, bmask, list, far[other], other,
                                   g->out.rowptr, ctr);
                GHET_HIP(hipGetLastError());
                GHET_HIP(hipMemcpyAsync(g->pinned, ctr, sizeof(MssCtr), hipMemcpyDeviceToHost, s));
                GHET_HIP(hipStreamSynchronize(s));
                F = hc->list_count;
                E = static_cast<int64_t>(slot_sum(hc->slots, 0));
                far_count = hc->far_count[other];
                unsigned long long mb = kInfBits;
                for (int i = 0; i < kSlots; ++i) mb = std::min(mb, hc->slots[i].v[1]);
                double far_min;
                std::memcpy(&far_min, &mb, sizeof(double));
                sel = other;
                ++rounds;
                if (F > 0 || far_count == 0) break;
                // nothing in the next bucket: jump to the smallest far distance
                old_thr = new_thr;
                new_thr = std::max(far_min + delta, std::nextafter(far_min, HUGE_VAL));
            }
            threshold = new_thr;
            list_len = F;
            list_edges = E;
            cur = list;
            nxt = qa;
            if (F == 0 && far_count == 0) break;
        }
        if (st) {
            hipLaunchKernelGGL(k_mss_reset, dim3(1), dim3(512), 0, s, ctr, -1);
            hipLaunchKernelGGL(k_mss_reached, dim3(2048), dim3(256), 0, s, dist, kn, kp, inf, g->out.rowptr, ctr->slots);
            GHET_HIP(hipMemcpyAsync(g->pinned, ctr, sizeof(MssCtr), hipMemcpyDeviceToHost, s));
        }
        if (dist_out) {
            hipLaunchKernelGGL(k_mss_transpose,
                               dim3(static_cast<unsigned>(ceil_div(n, 64)), static_cast<unsigned>(ceil_div(kp, 64))),
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
        st->extra[2] = static_cast<double>(scanned);  // edges scanned over all passes
    }
    return GHET_OK;
}

}  // namespace ghet
