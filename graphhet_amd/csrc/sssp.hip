// Single-source shortest paths in FP64 on gfx950: near-far bucketing with a
// light/heavy edge split.
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
// A bucket is the distance range below the threshold `thr` not yet settled.
// Each round expands the near queue edge-balanced (degree scan + 1024-edge
// tiles) but relaxes only the edges with d[u] + w < thr (light for the current
// d[u]); an improved vertex joins the next near queue.  Per-vertex round stamps
// keep the queue duplicate-free, and a vertex's first entry into a bucket also
// appends it to the bucket list.  When the near queue drains, every distance
// below thr is final, so one heavy pass over the bucket list relaxes the
// remaining edges (d[u] + w >= thr) exactly once and sends improved vertices
// to the far list (generation stamps keep it duplicate-free).  A vertex that
// re-enters the near queue k times inside a bucket thus gathers its heavy
// edges once instead of k times.  The far list is then split at thr + delta;
// the vertices below it form the next bucket list and its first near queue.
#include <cfloat>
#include <cmath>
#include <cstring>

#include "graph_internal.hpp"
#include "sched.hpp"

#include <rocprim/rocprim.hpp>

namespace ghet {
namespace {

// Queue positions (one atomic per LDS flush) plus spread sums:
// slot v[0] out-edges of the next near queue, v[1] bits of the smallest kept
// far distance (atomicMin on non-negative doubles' bit patterns), v[2]
// out-edges of the vertices appended to the bucket list, v[3] edges scanned
// (v[7] and done count block arrivals for the mailbox post).  list_count is the
// running length of the bucket lists (all buckets share one n-entry buffer:
// every vertex is listed in exactly one bucket) and is never reset.
struct SsspCtr {
    int32_t near_count;
    int32_t far_count[2];
    int32_t list_count;
    unsigned long long alloc[2];  // k_row_prep reservations, alternating by pass
    unsigned int done;            // blocks finished (mailbox ticket)
    int32_t pad[7];
    CtrSlot slots[kSlots];
};
constexpr int kCtrWords = sizeof(SsspCtr) / 8;
constexpr unsigned long long kInfBits = 0x7FF0000000000000ULL;

constexpr int kBT = 256;
constexpr int kItems = 4;
constexpr int kTile = kBT * kItems;  // edges per relax tile
// LDS capacity of each staged queue: a tile pushes at most kTile entries and
// a flush empties it above kOut - kTile.  1.25 kTile (instead of 2 kTile)
// lets 7 relax blocks share a CU instead of 5: s24 13.3 -> 12.3 ms.
constexpr int kOut = kTile + kTile / 4;
constexpr int kMaxBlocks = 2048;
constexpr int kSplitItems = 4;  // far-list entries per k_far_split thread and step
// A staged queue must hold what one step can push on top of what a flush
// leaves: a relax tile pushes at most kTile entries per queue (flush above
// kOut - kTile), a far-split step kBT * kSplitItems (flush above
// kOut - kBT * kSplitItems).
static_assert(kOut >= kTile && kOut >= kBT * kSplitItems, "staged queues smaller than one step's pushes");

__device__ __forceinline__ int upper_bound_i32(const int32_t* a, int len, int32_t key) {
    int lo = 0, hi = len;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (a[mid] <= key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// End of the light prefix of a weight-ordered row [b, e): the first i with
// du + w[i] >= thr (e if none).  Galloping from b (probes b, b+2, b+5,
// b+9, ...: gaps of 1, 2, 4, ...) and then a binary search inside the last
// gap.  Light prefixes are short next to their rows (about a bucket width
// out of weights spread over (0, 1]), and the first probes share a cache
// line, so most rows take one memory round trip instead of log2(degree).
__device__ __forceinline__ int32_t light_end(const double* __restrict__ w, int32_t b, int32_t e, double du,
                                             double thr) {
    int32_t lo = b, probe = b, gap = 1;  // every i < lo is light
    while (probe < e && du + w[probe] < thr) {
        lo = probe + 1;
        probe = lo + gap;
        gap <<= 1;
    }
    int32_t hi = probe < e ? probe : e;  // w[hi] is heavy, or hi == e
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (du + w[mid] < thr) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// The relax launches find each tile's entries through tile_first: entry
// tile_first[k] holds edge k * kTile of the pass (the preps write it: an
// entry with edges [a, a + c) marks every k with a <= k kTile < a + c), and
// tile_first[ceil(E / kTile)] = F - 1 closes the last tile.  A tile's
// entries are then tile_first[t] .. tile_first[t + 1] (the last may start
// past the tile: its edges are never looked up), with no binary search of
// the pass's scan per tile.
__device__ __forceinline__ void mark_tiles(int32_t* tile_first, int32_t slot, int32_t a, int32_t c) {
    for (int64_t k = (static_cast<int64_t>(a) + kTile - 1) / kTile; k * kTile < static_cast<int64_t>(a) + c; ++k)
        tile_first[k] = slot;
}
__device__ __forceinline__ void close_tiles(int32_t* tile_first, int32_t F, int32_t E) {
    tile_first[(static_cast<int64_t>(E) + kTile - 1) / kTile] = F - 1;
}

__global__ void k_sssp_seed(double* dist, int32_t* list, int32_t* stamp_near, int32_t src, SsspCtr* ctr) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        dist[src] = 0.0;
        list[0] = src;
        stamp_near[src] = 0;
        ctr->list_count = 1;
    }
}

// Resets the near queue and the spread sums; keeps the far counts (round) or
// also clears far list `other` (split).
__global__ void k_ctr_reset(SsspCtr* c, int other) {
    unsigned long long* w = reinterpret_cast<unsigned long long*>(c->slots);
    for (int i = threadIdx.x; i < kSlots * 8; i += blockDim.x) w[i] = (i % 8 == 1) ? kInfBits : 0ULL;
    if (threadIdx.x == 0) {
        c->near_count = 0;
        c->done = 0;
        if (other >= 0) c->far_count[other] = 0;
    }
}

// The pass's last block posts what the host loop reads: slot sums v0 (edges
// of the next near queue), v2 (edges of new list entries), v3 (edges
// scanned), the slot minimum v1 (far-list distance bits), and the near, far
// and list counts.  Block-uniform.
__device__ __forceinline__ void post_pass(SsspCtr* ctr, HostMail* mail, unsigned long long seq,
                                          unsigned long long token) {
    if (arrive_last(ctr->slots, &ctr->done, token) && threadIdx.x < 64) {  // wave 0: lane i reads slot i
        static_assert(kSlots == 64, "one slot per lane");
        const CtrSlot* sl = ctr->slots + threadIdx.x;
        const unsigned long long a = wave_sum(load_agent(&sl->v[0]));
        const unsigned long long mn = wave_min(load_agent(&sl->v[1]));
        const unsigned long long c = wave_sum(load_agent(&sl->v[2]));
        const unsigned long long d = wave_sum(load_agent(&sl->v[3]));
        if (threadIdx.x == 0) {
            auto ld32 = [](const int32_t* p) {
                return static_cast<unsigned long long>(__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            };
            mail->v[0] = a;
            mail->v[1] = mn;
            mail->v[2] = c;
            mail->v[3] = d;
            mail->v[4] = ld32(&ctr->near_count);
            mail->v[5] = ld32(&ctr->far_count[0]);
            mail->v[6] = ld32(&ctr->far_count[1]);
            mail->v[7] = ld32(&ctr->list_count);
            mail_post(mail, seq);
        }
    }
}

// Per queue entry u, the row segment a pass relaxes: rows are ordered by
// weight, so fl(d[u] + w) is non-decreasing along the row and the edges with
// d[u] + w < thr are a prefix (binary search).  Round: the prefix; heavy: the
// rest.  Two launch shapes prepare a pass: k_row_prep and k_row_split_small.
//
// k_row_prep, for queues of more than kSmallMax entries, in one launch: each
// block takes kPrepItems * kPrepBT queue entries (thread-major), splits their
// rows (light_end), scans the counts and
// reserves a range of entries and of edges with one 64-bit atomic on alloc
// ((entries << 32) | edges).  At its entry range it writes scan[] (first
// edge), start[] (first position) and vert[] (the vertex): blocks land in any
// order, but scan[] ascends, and the block that ends at F writes scan[F].
// Block 0 also resets the pass's counters (as k_ctr_reset with no far list;
// only the relax launch after this one uses them) and clears next_alloc, the
// other reservation counter, for the next pass.
constexpr int kPrepBT = 1024;  // entries per k_row_prep block: one reservation atomic each
constexpr int kPrepItems = 1;

template <bool kHeavy>
__global__ void __launch_bounds__(kPrepBT) k_row_prep(const int32_t* __restrict__ q, int32_t F,
                                                  const int32_t* __restrict__ rowptr, const double* __restrict__ w,
                                                  const double* __restrict__ dist, double threshold,
                                                  int32_t* __restrict__ scan, int32_t* __restrict__ start,
                                                  int32_t* __restrict__ vert, int32_t* __restrict__ tile_first,
                                                  unsigned long long* alloc, unsigned long long* next_alloc,
                                                  SsspCtr* ctr) {
    __shared__ int32_t s_wave[kPrepBT / 64];
    __shared__ unsigned long long s_base;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (blockIdx.x == 0) {
        unsigned long long* words = reinterpret_cast<unsigned long long*>(ctr->slots);
        for (int i = tid; i < kSlots * 8; i += kPrepBT) words[i] = (i % 8 == 1) ? kInfBits : 0ULL;
        if (tid == 0) {
            ctr->near_count = 0;
            ctr->done = 0;
            *next_alloc = 0;
        }
    }
    const int64_t i0 = static_cast<int64_t>(blockIdx.x) * kPrepBT * kPrepItems;
    int32_t u[kPrepItems], lo[kPrepItems], b[kPrepItems], e[kPrepItems];
#pragma unroll
    for (int k = 0; k < kPrepItems; ++k) {
        const int64_t i = i0 + static_cast<int64_t>(tid) * kPrepItems + k;
        u[k] = i < F ? q[i] : 0;
    }
#pragma unroll
    for (int k = 0; k < kPrepItems; ++k) {
        const int64_t i = i0 + static_cast<int64_t>(tid) * kPrepItems + k;
        b[k] = rowptr[u[k]];
        e[k] = i < F ? rowptr[u[k] + 1] : b[k];
        lo[k] = light_end(w, b[k], e[k], dist[u[k]], threshold);
    }
    int32_t c[kPrepItems];
    int32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kPrepItems; ++k) {
        c[k] = kHeavy ? e[k] - lo[k] : lo[k] - b[k];
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
    for (int v = 0; v < kPrepBT / 64; ++v) {
        if (v < wave) run += s_wave[v];
        total += s_wave[v];
    }
    const int32_t cnt = F - i0 < kPrepBT * kPrepItems ? static_cast<int32_t>(F - i0) : kPrepBT * kPrepItems;
    if (tid == 0)
        s_base = atomicAdd(alloc, (static_cast<unsigned long long>(cnt) << 32) | static_cast<uint32_t>(total));
    __syncthreads();
    const int32_t ebase = static_cast<int32_t>(s_base >> 32), xbase = static_cast<int32_t>(s_base & 0xFFFFFFFFu);
#pragma unroll
    for (int k = 0; k < kPrepItems; ++k) {
        const int64_t i = i0 + static_cast<int64_t>(tid) * kPrepItems + k;
        if (i < F) {
            const int32_t slot = ebase + tid * kPrepItems + k;
            scan[slot] = xbase + run;
            start[slot] = kHeavy ? lo[k] : b[k];
            vert[slot] = u[k];
            mark_tiles(tile_first, slot, xbase + run, c[k]);
            run += c[k];
        }
    }
    if (tid == 0 && ebase + cnt == F) {
        scan[F] = xbase + total;
        close_tiles(tile_first, F, xbase + total);
    }
}

// Small queues (at most kSmallMax entries): one 1024-thread block resets the
// pass's counters (as k_ctr_reset with no far list), splits the rows and
// scans the counts in queue order.
constexpr int kSmallBT = 1024;
constexpr int kSmallItems = 4;
constexpr int kSmallMax = kSmallBT * kSmallItems;

template <bool kHeavy>
__global__ void __launch_bounds__(kSmallBT) k_row_split_small(const int32_t* __restrict__ q, int32_t F,
                                                              const int32_t* __restrict__ rowptr,
                                                              const double* __restrict__ w,
                                                              const double* __restrict__ dist, double threshold,
                                                              int32_t* __restrict__ scan, int32_t* __restrict__ start,
                                                              int32_t* __restrict__ tile_first, SsspCtr* ctr) {
    __shared__ int32_t s_wave[kSmallBT / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    unsigned long long* words = reinterpret_cast<unsigned long long*>(ctr->slots);
    for (int i = tid; i < kSlots * 8; i += kSmallBT) words[i] = (i % 8 == 1) ? kInfBits : 0ULL;
    if (tid == 0) {
        ctr->near_count = 0;
        ctr->done = 0;
        ctr->alloc[0] = 0;  // whichever k_row_prep runs next starts from zero
        ctr->alloc[1] = 0;
    }
    int32_t c[kSmallItems], st[kSmallItems];
    int32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kSmallItems; ++k) {
        const int i = tid * kSmallItems + k;
        c[k] = 0;
        st[k] = 0;
        if (i < F) {
            const int32_t u = q[i];
            const int32_t b = rowptr[u], e = rowptr[u + 1];
            const int32_t lo = light_end(w, b, e, dist[u], threshold);
            c[k] = kHeavy ? e - lo : lo - b;
            st[k] = kHeavy ? lo : b;
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
            mark_tiles(tile_first, i, run, c[k]);
            run += c[k];
        }
    }
    if (tid == 0) {
        scan[F] = total;
        close_tiles(tile_first, F, total);
    }
}

// One pass, persistent blocks over kTile-edge tiles of queue q; entry j
// covers edges [start[j], start[j] + scan[j+1] - scan[j]) and scan[F] is the
// pass's edge count (E_max, from the host, only bounds the grid).
// Round (kHeavy false): relaxes the edges with d[u] + w < thr; an improved v
// joins next_q (stamp_near != next_round) and, on its first entry into this
// bucket (old stamp < r0, the bucket's first round), the bucket list.
// Heavy (kRelaxHeavy): q is the drained bucket's list; relaxes the edges with
// d[u] + w >= thr and sends improved vertices to far list far_sel.
// Pull (kRelaxPull): the heavy pass from the other side.  q holds the
// unsettled vertices v (d[v] >= thr) with in-edges, rowptr/col/wts are the
// in-CSR, and an edge (x, v) relaxes v from x when x is settled (d[x] < thr).
// Edges from earlier buckets' vertices change nothing (their heavy passes
// already relaxed them), and no settled x has d[x] + w < thr into an
// unsettled v (the rounds relaxed those), so v ends at the same minimum as
// after the push pass and joins the far list under the same stamp rule.
// Targets stay at or above thr, sources below: no edge reads a distance
// another edge writes.
enum RelaxMode { kRelaxRound, kRelaxHeavy, kRelaxPull };

template <int kMode>
__global__ void __launch_bounds__(kBT) k_relax(const int32_t* __restrict__ q, int32_t F,
                                               const int32_t* __restrict__ scan,
                                               const int32_t* __restrict__ start,
                                               const int32_t* __restrict__ tile_first,
                                               const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                               const double* __restrict__ wts, double* dist, double threshold,
                                               int32_t* stamp_near, int32_t next_round, int32_t r0,
                                               int32_t* __restrict__ next_q, int32_t* __restrict__ list,
                                               int32_t* stamp_far, int32_t far_gen, int32_t* far_list, int far_sel,
                                               SsspCtr* ctr, HostMail* mail, unsigned long long seq) {
    __shared__ int32_t s_scan[kTile];
    __shared__ int32_t s_start[kTile];
    __shared__ int32_t s_src[kTile];
    __shared__ int32_t s_qa[kOut], s_qb[kOut];  // round: near, list; heavy: far
    __shared__ int32_t s_i0, s_span, s_na, s_nb, s_base;
    __shared__ unsigned long long s_red[kBT / 64];
    constexpr bool kHeavy = kMode != kRelaxRound;  // heavy and pull feed the far list
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
    unsigned long long token = 0;  // thread 0: results of its counter atomics (post_pass)
    if (blockIdx.x == 0 && tid == 0) token ^= slot_add_ret(ctr->slots, 3, static_cast<unsigned long long>(E));
    for (int64_t tile = blockIdx.x; tile * kTile < E; tile += gridDim.x) {
        const int32_t t0 = static_cast<int32_t>(tile * kTile);
        const int32_t t1 = static_cast<int32_t>(std::min<int64_t>(static_cast<int64_t>(t0) + kTile, E));
        if (tid == 0) {
            const int i0 = tile_first[tile];      // holds edge t0
            const int i1 = tile_first[tile + 1];  // holds edge t1, or the pass's last entry
            s_i0 = i0;
            s_span = i1 - i0 + 1;
        }
        __syncthreads();
        const int i0 = s_i0, span = s_span;
        const bool use_lds = span <= kTile;
        if (use_lds) {
            for (int j = tid; j < span; j += kBT) {
                s_scan[j] = scan[i0 + j];
                s_start[j] = start[i0 + j];
                s_src[j] = q[i0 + j];
            }
        }
        __syncthreads();
        // The thread's kItems edges go through each stage together, so their
        // loads (row distance, weight, target, then the target distances)
        // are in flight at the same time instead of one edge after another.
        int32_t u[kItems], pos[kItems];
        bool ok[kItems];
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            const int32_t e = t0 + k * kBT + tid;
            ok[k] = e < t1;
            u[k] = 0;
            pos[k] = 0;  // E > 0, so edge 0 exists: masked lanes load it
            if (ok[k]) {
                if (use_lds) {
                    const int j = upper_bound_i32(s_scan, span, e) - 1;
                    pos[k] = s_start[j] + (e - s_scan[j]);
                    u[k] = s_src[j];
                } else {
                    const int j = upper_bound_i32(scan + i0, span, e) - 1;
                    u[k] = q[i0 + j];
                    pos[k] = start[i0 + j] + (e - scan[i0 + j]);
                }
            }
        }
        double nd[kItems];
        int32_t v[kItems];
        bool cand[kItems];
        double dv[kItems];
        if (kMode == kRelaxPull) {
            // u[k] is the target; its in-neighbour, the weight and the
            // target's distance load together, then the neighbour's distance
            int32_t x[kItems];
            double wv[kItems];
#pragma unroll
            for (int k = 0; k < kItems; ++k) {
                x[k] = col[pos[k]];
                wv[k] = wts[pos[k]];
                dv[k] = dist[u[k]];
                v[k] = u[k];
            }
#pragma unroll
            for (int k = 0; k < kItems; ++k) {
                const double dx = dist[x[k]];
                nd[k] = dx + wv[k];
                cand[k] = ok[k] && dx < threshold;
            }
        } else {
#pragma unroll
            for (int k = 0; k < kItems; ++k) {
                nd[k] = dist[u[k]] + wts[pos[k]];
                v[k] = col[pos[k]];
            }
#pragma unroll
            for (int k = 0; k < kItems; ++k) {
                cand[k] = ok[k] && (kHeavy ? nd[k] >= threshold : nd[k] < threshold);
                dv[k] = cand[k] ? dist[v[k]] : 0.0;
            }
        }
        if (kMode == kRelaxPull) {
            // The lanes of one item hold consecutive edges, and each target
            // is one entry, so its edges are a run of lanes: a segmented
            // minimum leaves the run's smallest candidate in the run's last
            // lane, which alone updates the target (one atomic per run
            // instead of one per edge).
            const int lane = tid & 63;
#pragma unroll
            for (int k = 0; k < kItems; ++k) {
                double x = cand[k] ? nd[k] : HUGE_VAL;
                for (int off = 1; off < 64; off <<= 1) {
                    const double y = __shfl_up(x, off);
                    const int32_t yv = __shfl_up(v[k], off);
                    if (lane >= off && yv == v[k] && y < x) x = y;
                }
                const int32_t nv = __shfl_down(v[k], 1);
                nd[k] = x;
                cand[k] = (lane == 63 || nv != v[k]) && x < HUGE_VAL;
            }
        }
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            bool push_a = false, push_b = false;
            if (cand[k] && nd[k] < dv[k]) {
                const double old = atomic_min_nonneg(&dist[v[k]], nd[k]);
                if (nd[k] < old) {
                    if (kHeavy) {
                        push_a = atomicExch(&stamp_far[v[k]], far_gen) != far_gen;
                    } else {
                        const int32_t prev = atomicExch(&stamp_near[v[k]], next_round);
                        push_a = prev != next_round;
                        push_b = prev < r0;
                    }
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
            token ^= slot_add_ret(ctr->slots, 0, od);
            token ^= slot_add_ret(ctr->slots, 2, ld);
        }
    }
    post_pass(ctr, mail, seq, token);
}

// Preparation of a pull pass: the unsettled vertices (d >= thr) with
// in-edges, as entries in the format k_relax reads (scan[] first edge,
// start[] first in-CSR position, vert[] the vertex).  Persistent blocks
// (about two per CU) each own a run of kPullChunk-vertex chunks.  A block
// first counts its run's entries and in-edges, reserves both ranges with ONE
// 64-bit atomic on alloc ((entries << 32) | edges; blocks without entries
// skip it), then reads the run again (mostly from the caches), scans each
// chunk (item-major: every load is coalesced) and writes its entries (at
// R-MAT scale 24: 115 us, against 270 us with a block and a reservation per
// chunk and one read of the vertices).  The last block to arrive writes
// scan[F] = E, posts E (v[0]), F (v[4]) and the far and list counts, and
// clears the counters for the relax launch that follows.  The counters must
// be zero before the launch; next_alloc, the other reservation word, is
// cleared for the next pass.
constexpr int kPullBT = 1024;
constexpr int kPullItems = 4;
constexpr int kPullChunk = kPullBT * kPullItems;

__global__ void __launch_bounds__(kPullBT) k_pull_prep(int32_t n, const int32_t* __restrict__ in_rowptr,
                                                      const double* __restrict__ dist, double threshold,
                                                      int32_t* __restrict__ scan, int32_t* __restrict__ start,
                                                      int32_t* __restrict__ vert, int32_t* __restrict__ tile_first,
                                                      unsigned long long* alloc, unsigned long long* next_alloc,
                                                      SsspCtr* ctr, HostMail* mail, unsigned long long seq) {
    constexpr int W = kPullBT / 64;
    __shared__ int32_t s_ent[kPullItems][W], s_edg[kPullItems][W];
    __shared__ unsigned long long s_red[W];
    __shared__ unsigned long long s_base;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (blockIdx.x == 0 && tid == 0) *next_alloc = 0;
    const int64_t chunks = ceil_div(n, kPullChunk);
    const int64_t c0 = chunks * blockIdx.x / gridDim.x, c1 = chunks * (blockIdx.x + 1) / gridDim.x;
    // in-edges of vertex v if it is unsettled, else 0; n > 0, so lanes past
    // the end read vertex 0 and count nothing
    auto load = [&](int64_t c, int32_t* b, int32_t* cnt) {
        double d[kPullItems];
        int32_t e[kPullItems];
#pragma unroll
        for (int k = 0; k < kPullItems; ++k) {
            const int64_t v = c * kPullChunk + k * kPullBT + tid;
            const int64_t vv = v < n ? v : 0;
            d[k] = dist[vv];
            b[k] = in_rowptr[vv];
            e[k] = in_rowptr[vv + 1];
            if (v >= n) e[k] = b[k];
        }
#pragma unroll
        for (int k = 0; k < kPullItems; ++k) cnt[k] = d[k] >= threshold ? e[k] - b[k] : 0;
    };
    // pass 1: the run's totals, packed (entries << 32) + edges (the edges of
    // all vertices are below 2^31, so the low half never carries)
    unsigned long long mine = 0;
    for (int64_t c = c0; c < c1; ++c) {
        int32_t b[kPullItems], cnt[kPullItems];
        load(c, b, cnt);
#pragma unroll
        for (int k = 0; k < kPullItems; ++k)
            mine += (cnt[k] > 0 ? (1ULL << 32) : 0ULL) + static_cast<unsigned long long>(cnt[k]);
    }
    const unsigned long long tot = block_sum_u64<kPullBT>(mine, s_red);
    if (tid == 0 && tot != 0) s_base = atomicAdd(alloc, tot);
    __syncthreads();
    unsigned long long token = 0;
    if (tot != 0) {
        const unsigned long long sb = s_base;
        if (tid == 0) token = sb;
        int32_t ebase = static_cast<int32_t>(sb >> 32), xbase = static_cast<int32_t>(sb & 0xFFFFFFFFu);
        const unsigned long long below = (1ULL << lane) - 1;
        // pass 2: chunk by chunk, entries in (item, wave, lane) order
        for (int64_t c = c0; c < c1; ++c) {
            int32_t b[kPullItems], cnt[kPullItems], incl[kPullItems], rank[kPullItems];
            load(c, b, cnt);
#pragma unroll
            for (int k = 0; k < kPullItems; ++k) {
                incl[k] = cnt[k];
                for (int off = 1; off < 64; off <<= 1) {
                    const int32_t y = __shfl_up(incl[k], off);
                    if (lane >= off) incl[k] += y;
                }
                const unsigned long long mask = __ballot(cnt[k] > 0);
                rank[k] = __popcll(mask & below);
                if (lane == 63) s_edg[k][wave] = incl[k];
                if (lane == 0) s_ent[k][wave] = __popcll(mask);
            }
            __syncthreads();
            int32_t ent_off[kPullItems] = {}, edg_off[kPullItems] = {};
            int32_t ent_tot = 0, edg_tot = 0;  // block-uniform
#pragma unroll
            for (int k = 0; k < kPullItems; ++k)
                for (int w = 0; w < W; ++w) {
                    if (w == wave) {
                        ent_off[k] = ent_tot;
                        edg_off[k] = edg_tot;
                    }
                    ent_tot += s_ent[k][w];
                    edg_tot += s_edg[k][w];
                }
            __syncthreads();  // the next chunk rewrites s_ent / s_edg
#pragma unroll
            for (int k = 0; k < kPullItems; ++k)
                if (cnt[k] > 0) {
                    const int32_t slot = ebase + ent_off[k] + rank[k];
                    const int32_t first_edge = xbase + edg_off[k] + incl[k] - cnt[k];
                    scan[slot] = first_edge;
                    start[slot] = b[k];
                    vert[slot] = static_cast<int32_t>(c * kPullChunk + k * kPullBT + tid);
                    mark_tiles(tile_first, slot, first_edge, cnt[k]);
                }
            ebase += ent_tot;
            xbase += edg_tot;
        }
    }
    if (arrive_last(ctr->slots, &ctr->done, token)) {
        if (tid == 0) {
            auto ld32 = [](const int32_t* p) {
                return static_cast<unsigned long long>(__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            };
            const unsigned long long a = load_agent(alloc);
            const int32_t F = static_cast<int32_t>(a >> 32), E = static_cast<int32_t>(a & 0xFFFFFFFFu);
            scan[F] = E;
            close_tiles(tile_first, F, E);
            mail->v[0] = static_cast<unsigned long long>(E);
            mail->v[1] = kInfBits;
            mail->v[2] = 0;
            mail->v[3] = 0;
            mail->v[4] = static_cast<unsigned long long>(F);
            mail->v[5] = ld32(&ctr->far_count[0]);
            mail->v[6] = ld32(&ctr->far_count[1]);
            mail->v[7] = ld32(&ctr->list_count);
            mail_post(mail, seq);
        }
        // every other block has arrived: nothing reads the counters again
        unsigned long long* words = reinterpret_cast<unsigned long long*>(ctr->slots);
        for (int i = tid; i < kSlots * 8; i += kPullBT) words[i] = (i % 8 == 1) ? kInfBits : 0ULL;
        if (tid == 0) {
            ctr->near_count = 0;
            ctr->done = 0;
        }
    }
}

// Splits far list `sel` at new_thr: entries below old_thr are stale (already
// settled at their current distance), [old_thr, new_thr) are appended to the
// bucket list (which is also the new bucket's first near queue), the rest go
// to far list 1-sel with the new generation stamp.
__global__ void __launch_bounds__(kBT) k_far_split(const int32_t* __restrict__ far_in, int32_t count,
                                                   const double* __restrict__ dist, double old_thr, double new_thr,
                                                   int32_t* stamp_near, int32_t next_round, int32_t* stamp_far,
                                                   int32_t new_gen, int32_t* __restrict__ list, int32_t* far_out,
                                                   int other, const int32_t* __restrict__ rowptr, SsspCtr* ctr,
                                                   HostMail* mail, unsigned long long seq) {
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
    // kSplitItems entries per thread per step: their vertex and distance
    // loads are in flight together
    for (int64_t base = static_cast<int64_t>(blockIdx.x) * kBT * kSplitItems; base < count;
         base += static_cast<int64_t>(gridDim.x) * kBT * kSplitItems) {
        int32_t v[kSplitItems];
        double d[kSplitItems];
#pragma unroll
        for (int k = 0; k < kSplitItems; ++k) {
            const int64_t i = base + k * kBT + threadIdx.x;
            v[k] = i < count ? far_in[i] : -1;
        }
#pragma unroll
        for (int k = 0; k < kSplitItems; ++k) d[k] = v[k] >= 0 ? dist[v[k]] : 0.0;
#pragma unroll
        for (int k = 0; k < kSplitItems; ++k) {
            bool to_near = false, keep = false;
            if (v[k] >= 0 && d[k] >= old_thr) {
                if (d[k] < new_thr) {
                    stamp_near[v[k]] = next_round;
                    to_near = true;
                } else {
                    stamp_far[v[k]] = new_gen;
                    keep = true;
                    const unsigned long long bits = static_cast<unsigned long long>(__double_as_longlong(d[k]));
                    dmin = bits < dmin ? bits : dmin;
                }
            }
            lq_push(s_near, &s_nc, to_near, v[k]);
            lq_push(s_far, &s_fc, keep, v[k]);
        }
        lq_flush<kBT>(s_near, &s_nc, &s_base, &ctr->list_count, list, kOut - kBT * kSplitItems, on_near);
        lq_flush<kBT>(s_far, &s_fc, &s_base, &ctr->far_count[other], far_out, kOut - kBT * kSplitItems, nop);
    }
    lq_flush<kBT>(s_near, &s_nc, &s_base, &ctr->list_count, list, 0, on_near);
    lq_flush<kBT>(s_far, &s_fc, &s_base, &ctr->far_count[other], far_out, 0, nop);
    od = block_sum_u64<kBT>(od, s_red);
    dmin = wave_min(dmin);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = dmin;
    __syncthreads();
    unsigned long long token = 0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < kBT / 64; ++w) dmin = s_red[w] < dmin ? s_red[w] : dmin;
        token ^= slot_add_ret(ctr->slots, 0, od);
        if (dmin != kInfBits) token ^= atomicMin(&ctr->slots[blockIdx.x & (kSlots - 1)].v[1], dmin);
    }
    post_pass(ctr, mail, seq, token);
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

// Sort keys of the weights: the bit patterns of non-negative doubles order as
// unsigned integers (-0.0 is mapped to +0.0).
__global__ void k_weight_keys(const double* __restrict__ w, int64_t m, unsigned long long* __restrict__ key) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < m;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const double x = w[i];
        key[i] = x == 0.0 ? 0ULL : static_cast<unsigned long long>(__double_as_longlong(x));
    }
}

// Orders each out-row by weight into g->sl, built once per graph on the first
// SSSP: a stable segmented radix sort (rocPRIM) of (weight bits, target) with
// one segment per row, so equal weights keep the row's target order
// (ascending in a canonical CSR).  The sorted keys are the weights.
int build_sssp_layout(ghet_graph* g) {
    SsspLayout& L = g->sl;
    if (L.built) return GHET_OK;
    const int64_t m = g->out.m;
    const int32_t n = g->n;
    if (m > 0xFFFFFFFFLL) return set_error(GHET_ERR_UNSUPPORTED, "SSSP layout: %lld edges", static_cast<long long>(m));
    const size_t col_bytes = (4 * size_t(m) + 255) & ~size_t(255);
    GHET_HIP(hipMalloc(&L.mem, col_bytes + 8 * size_t(std::max<int64_t>(m, 1))));
    L.col = static_cast<int32_t*>(L.mem);
    L.w = reinterpret_cast<double*>(static_cast<char*>(L.mem) + col_bytes);
    if (m > 0) {
        hipStream_t s = g->stream;
        unsigned long long* keys = nullptr;
        void* tmp = nullptr;
        size_t tb = 0;
        auto sort = [&](void* t) {
            return rocprim::segmented_radix_sort_pairs(
                t, tb, keys, reinterpret_cast<unsigned long long*>(L.w), g->out.col, L.col, static_cast<unsigned>(m),
                static_cast<unsigned>(n), g->out.rowptr, g->out.rowptr + 1, 0, 64, s);
        };
        hipError_t e = hipMalloc(&keys, 8 * size_t(m));
        if (e == hipSuccess) {
            const unsigned grid = static_cast<unsigned>(std::min<int64_t>(ceil_div(m, 256), 65536));
            hipLaunchKernelGGL(k_weight_keys, dim3(grid), dim3(256), 0, s, g->out.w, m, keys);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = sort(nullptr);
        if (e == hipSuccess) e = hipMalloc(&tmp, std::max<size_t>(tb, 256));
        if (e == hipSuccess) e = sort(tmp);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (tmp) hipFree(tmp);
        if (keys) hipFree(keys);
        if (e != hipSuccess) {
            hipFree(L.mem);
            L = SsspLayout{};
            (void)hipGetLastError();
            return set_error(GHET_ERR_HIP, "SSSP layout sort: %s", hipGetErrorString(e));
        }
    }
    L.built = true;
    return GHET_OK;
}

__global__ void __launch_bounds__(256) k_wsum(const double* __restrict__ w, int64_t m, double* part) {
    __shared__ double s_red[4];
    double x = 0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < m; i += static_cast<int64_t>(gridDim.x) * 256)
        x += w[i];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = s_red[0] + s_red[1] + s_red[2] + s_red[3];
}

// Base bucket width: 4 mean weights per mean out-degree, so a bucket spans
// about four light edges of an average row; light/heavy splitting keeps the
// rest of each row out of the bucket's rounds.  With size_term, graphs below
// 2^24 vertices get wider buckets, x2^(0.8 (24 - log2 n)) up to x64: their
// buckets hold too little work to pay a pass's fixed cost (fitted to the
// best measured widths on R-MAT scales 18-24).  The autotuner's class scales
// it by 2^(class-1).  The mean weight is reduced once per graph.
int sssp_base_delta(ghet_graph* g, double* out, bool size_term) {
    if (g->mean_w < 0) {
        const int64_t m = g->out.m;
        double mean = 1.0;
        if (m > 0 && g->out.w) {
            constexpr int kParts = 1024;
            double* part;
            GHET_HIP(hipMalloc(&part, sizeof(double) * kParts));
            hipLaunchKernelGGL(k_wsum, dim3(kParts), dim3(256), 0, g->stream, g->out.w, m, part);
            std::vector<double> hp(kParts);
            const hipError_t e1 = hipMemcpyAsync(hp.data(), part, sizeof(double) * kParts, hipMemcpyDeviceToHost, g->stream);
            const hipError_t e2 = hipStreamSynchronize(g->stream);
            hipFree(part);
            GHET_HIP(e1);
            GHET_HIP(e2);
            double sum = 0;
            for (double x : hp) sum += x;
            mean = sum / static_cast<double>(m);
        }
        g->mean_w = mean > 0 ? mean : 1.0;
    }
    const double avg_deg = g->n > 0 ? static_cast<double>(g->out.m) / g->n : 1.0;
    const double small =
        size_term ? std::min(6.0, std::max(0.0, 0.8 * (24.0 - std::log2(std::max<double>(g->n, 1))))) : 0.0;
    *out = std::max(1e-12, 4.0 * g->mean_w / std::max(avg_deg, 1.0)) * std::exp2(small) *
           std::ldexp(1.0, g->tune.sssp_class - 1);
    return GHET_OK;
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
    double base_delta = 0;
    if (have_opts && o->sssp_delta > 0) base_delta = o->sssp_delta;
    else GHET_CHECK(sssp_base_delta(g, &base_delta, true));
    ghet_level_trace* trace = have_opts ? o->trace : nullptr;
    const int trace_cap = trace ? o->trace_cap : 0;
    int trace_len = 0;

    double* dist;
    int32_t *qa, *qb, *scan, *qvert, *stamp_near, *stamp_far, *far[2], *list, *qstart;
    SsspCtr* ctr;
    void* p;
    GHET_CHECK(g->ws.get(WS_DIST, 8 * size_t(n), &p)); dist = static_cast<double*>(p);
    GHET_CHECK(g->ws.get(WS_QUEUE_A, 4 * size_t(n), &p)); qa = static_cast<int32_t*>(p);
    GHET_CHECK(g->ws.get(WS_QUEUE_B, 4 * size_t(n), &p)); qb = static_cast<int32_t*>(p);
    GHET_CHECK(g->ws.get(WS_FAR_A, 4 * size_t(n), &p)); far[0] = static_cast<int32_t*>(p);
    GHET_CHECK(g->ws.get(WS_FAR_B, 4 * size_t(n), &p)); far[1] = static_cast<int32_t*>(p);
    GHET_CHECK(g->ws.get(WS_SSSP_LIST, 4 * size_t(n), &p)); list = static_cast<int32_t*>(p);
    GHET_CHECK(g->ws.get(WS_SSSP_START, 4 * size_t(n), &p)); qstart = static_cast<int32_t*>(p);
    GHET_CHECK(g->ws.get(WS_STAMP_NEAR, 4 * size_t(n), &p)); stamp_near = static_cast<int32_t*>(p);
    GHET_CHECK(g->ws.get(WS_STAMP_FAR, 4 * size_t(n), &p)); stamp_far = static_cast<int32_t*>(p);
    GHET_CHECK(g->ws.get(WS_SCAN, 4 * (size_t(n) + 1), &p)); scan = static_cast<int32_t*>(p);
    GHET_CHECK(g->ws.get(WS_SSSP_VERT, 4 * size_t(n), &p)); qvert = static_cast<int32_t*>(p);
    GHET_CHECK(g->ws.get(WS_COUNTERS, sizeof(SsspCtr), &p)); ctr = static_cast<SsspCtr*>(p);
    // a pass's edges are distinct out-edges (or in-edges): at most m of them
    const int64_t m_pass = std::max<int64_t>(g->out.m, g->in.present() ? g->in.m : 0);
    GHET_CHECK(g->ws.get(WS_SSSP_TILE, 4 * (size_t(ceil_div(m_pass, kTile)) + 2), &p));
    int32_t* tile_first = static_cast<int32_t*>(p);
    GHET_CHECK(mail_init(g));
    const HostMail* hm = g->mail;  // totals of the last pass (mail_wait)

    GHET_CHECK(build_sssp_layout(g));
    SchedState sched;
    sched_begin(&sched, schedule, GHET_NET_SSSP, n, g->out.m, g->out.m);

    GHET_HIP(hipEventRecord(g->ev_begin, s));
    GHET_CHECK(fill<double>(dist, inf, n, s));
    GHET_CHECK(fill<int32_t>(stamp_near, -1, n, s));
    GHET_CHECK(fill<int32_t>(stamp_far, -1, n, s));
    GHET_CHECK(fill<uint64_t>(reinterpret_cast<uint64_t*>(ctr), 0, kCtrWords, s));
    hipLaunchKernelGGL(k_sssp_seed, dim3(1), dim3(64), 0, s, dist, list, stamp_near, src, ctr);

    // The near queue of a bucket's first round is its list segment; later
    // rounds alternate between qa and qb.
    const int32_t* cur = list;
    int32_t* nxt = qa;
    int32_t F = 1;
    int64_t E = g->h_out_rowptr[src + 1] - g->h_out_rowptr[src];
    int32_t list_base = 0, list_len = 1, list_count = 1;  // the current bucket's segment, all segments
    int64_t list_edges = E;
    int32_t round = 0, r0 = 0, far_gen = 0, phases = 0;
    int sel = 0;
    int32_t far_count = 0;
    double threshold = 0.0;
    double delta = base_delta;
    // Bucket-width decision: before the first bucket and at every threshold
    // advance.  A trace record keeps the features, the class, the threshold
    // and the device time so far (one event sync per decision, only when
    // tracing).
    auto decide = [&](LevelInfo& li) -> int {
        li.prev_mode = sched.sssp_prev;
        const int cls = sched_pick_sssp(&sched, li);
        if (trace && trace_len < trace_cap) {
            ghet_level_trace& t = trace[trace_len++];
            sched_features(li, t.features);
            t.mode = cls;
            t.threshold = li.threshold;
            float ms = 0;
            if (hipEventRecord(g->ev_k1, s) == hipSuccess && hipEventSynchronize(g->ev_k1) == hipSuccess)
                (void)hipEventElapsedTime(&ms, g->ev_begin, g->ev_k1);
            t.ms = ms;
        }
        return cls;
    };
    {
        LevelInfo li;
        li.n = n; li.m = g->out.m; li.m_in = g->out.m; li.level = 0; li.frontier = F; li.frontier_edges = E;
        li.unvisited = n - 1; li.delta = base_delta; li.threshold = 0; li.far_size = 0;
        const int cls = decide(li);
        delta = base_delta * std::ldexp(1.0, cls - 1);
        threshold = delta;
    }
    // Edge-balanced pass over queue q of F vertices whose rows hold at most
    // E_max edges in all (the grid bound); the pass covers each row's light
    // prefix (round) or heavy rest (heavy pass).
    const SsspLayout& L = g->sl;
    // The pass also resets the counters it fills.
    int64_t passes = 0;
    int cus = 0;  // the pull preparation's grid: two 1024-thread blocks per CU
    GHET_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device));
    cus = std::max(cus, 1);
    auto launch_pass = [&](bool heavy, const int32_t* q, int32_t Fq, int64_t E_max, int32_t next_round) -> int {
        if (Fq <= kSmallMax) {
            if (heavy)
                hipLaunchKernelGGL(k_row_split_small<true>, dim3(1), dim3(kSmallBT), 0, s, q, Fq, g->out.rowptr, L.w,
                                   dist, threshold, scan, qstart, tile_first, ctr);
            else
                hipLaunchKernelGGL(k_row_split_small<false>, dim3(1), dim3(kSmallBT), 0, s, q, Fq, g->out.rowptr, L.w,
                                   dist, threshold, scan, qstart, tile_first, ctr);
        } else {
            const unsigned gq = static_cast<unsigned>(ceil_div(Fq, kPrepBT * kPrepItems));
            unsigned long long* al = ctr->alloc + (passes & 1);
            unsigned long long* next_al = ctr->alloc + ((passes + 1) & 1);
            if (heavy)
                hipLaunchKernelGGL(k_row_prep<true>, dim3(gq), dim3(kPrepBT), 0, s, q, Fq, g->out.rowptr, L.w, dist,
                                   threshold, scan, qstart, qvert, tile_first, al, next_al, ctr);
            else
                hipLaunchKernelGGL(k_row_prep<false>, dim3(gq), dim3(kPrepBT), 0, s, q, Fq, g->out.rowptr, L.w, dist,
                                   threshold, scan, qstart, qvert, tile_first, al, next_al, ctr);
            q = qvert;  // the relax reads the entries in their reserved order
        }
        ++passes;
        const unsigned grid = static_cast<unsigned>(std::min<int64_t>(ceil_div(E_max, kTile), kMaxBlocks));
        const unsigned long long seq = mail_next(g);
        if (heavy)
            hipLaunchKernelGGL(k_relax<kRelaxHeavy>, dim3(grid), dim3(kBT), 0, s, q, Fq, scan, qstart, tile_first,
                               g->out.rowptr, L.col,
                               L.w, dist, threshold, stamp_near, next_round, r0, nxt, list, stamp_far, far_gen,
                               far[sel], sel, ctr, g->mail_dev, seq);
        else
            hipLaunchKernelGGL(k_relax<kRelaxRound>, dim3(grid), dim3(kBT), 0, s, q, Fq, scan, qstart, tile_first,
                               g->out.rowptr, L.col,
                               L.w, dist, threshold, stamp_near, next_round, r0, nxt, list, stamp_far, far_gen,
                               far[sel], sel, ctr, g->mail_dev, seq);
        GHET_HIP(hipGetLastError());
        GHET_CHECK(mail_wait(g, s, seq));
        return GHET_OK;
    };
    // Pull form of the heavy pass (k_pull_prep + k_relax<kRelaxPull>, over
    // the weighted in-CSR): once most of the graph is settled, the in-edges
    // of the unsettled vertices are far fewer than the heavy edges of a big
    // bucket.  Their count is estimated first as the out-edges of the
    // vertices not yet settled (exact when the graph is undirected); if that
    // is under half the bucket's out-edges, the preparation counts them
    // exactly, and the pull runs when they are fewer than the bucket's
    // out-edges.  Below n/4 bucket edges the preparation's scan of all
    // vertices costs more than it could save.  opts reserved[0] bit 3 keeps
    // every heavy pass a push, bit 4 pulls every heavy pass (tests).
    const int32_t heavy_flags = have_opts ? o->reserved[0] : 0;
    const bool can_pull = g->in.present() && g->in.w != nullptr && !(heavy_flags & 8);
    const bool force_pull = can_pull && (heavy_flags & 16);
    int64_t settled_out = 0;  // out-edges of the vertices of finished buckets
    int32_t pulls = 0;
    int64_t rounds = 0, relaxed = 0;
    // *pulled: the heavy pass ran as a pull.
    auto try_pull = [&](bool* pulled) -> int {
        *pulled = false;
        const int64_t est = g->out.m - settled_out - list_edges;
        if (!force_pull && !(list_edges >= n / 4 && est < list_edges / 2)) return GHET_OK;
        hipLaunchKernelGGL(k_ctr_reset, dim3(1), dim3(512), 0, s, ctr, -1);
        unsigned long long* al = ctr->alloc + (passes & 1);
        unsigned long long* next_al = ctr->alloc + ((passes + 1) & 1);
        ++passes;
        const unsigned gp = static_cast<unsigned>(std::min<int64_t>(ceil_div(n, kPullChunk), 2 * cus));
        unsigned long long seq = mail_next(g);
        hipLaunchKernelGGL(k_pull_prep, dim3(gp), dim3(kPullBT), 0, s, n, g->in.rowptr, dist, threshold, scan, qstart,
                           qvert, tile_first, al, next_al, ctr, g->mail_dev, seq);
        GHET_HIP(hipGetLastError());
        GHET_CHECK(mail_wait(g, s, seq));
        const int32_t Fp = static_cast<int32_t>(hm->v[4]);
        const int64_t Ep = static_cast<int64_t>(hm->v[0]);
        if (!force_pull && Ep >= list_edges) return GHET_OK;
        *pulled = true;
        ++pulls;
        if (Ep > 0) {
            const unsigned grid = static_cast<unsigned>(std::min<int64_t>(ceil_div(Ep, kTile), kMaxBlocks));
            seq = mail_next(g);
            hipLaunchKernelGGL(k_relax<kRelaxPull>, dim3(grid), dim3(kBT), 0, s, qvert, Fp, scan, qstart, tile_first,
                               g->in.rowptr, g->in.col, g->in.w, dist, threshold, stamp_near, round, r0, nxt, list,
                               stamp_far, far_gen, far[sel], sel, ctr, g->mail_dev, seq);
            GHET_HIP(hipGetLastError());
            GHET_CHECK(mail_wait(g, s, seq));
            relaxed += static_cast<int64_t>(hm->v[3]);
        }
        far_count = static_cast<int32_t>(hm->v[5 + sel]);
        return GHET_OK;
    };
    for (;;) {
        // relaxation rounds of the current bucket (light edges only)
        while (F > 0) {
            const int32_t next_round = round + 1;
            if (E > 0) {
                GHET_CHECK(launch_pass(false, cur, F, E, next_round));
                relaxed += static_cast<int64_t>(hm->v[3]);
                F = static_cast<int32_t>(hm->v[4]);
                E = static_cast<int64_t>(hm->v[0]);
                list_edges += static_cast<int64_t>(hm->v[2]);
                list_count = static_cast<int32_t>(hm->v[7]);
                list_len = list_count - list_base;
            } else {
                F = 0;
            }
            cur = nxt;
            nxt = (nxt == qa) ? qb : qa;
            round = next_round;
            ++rounds;
        }
        // heavy pass: the bucket's distances are final; its remaining edges
        // all land at or above the threshold
        if (list_edges > 0) {
            bool pulled = false;
            if (can_pull) GHET_CHECK(try_pull(&pulled));
            if (!pulled) {
                GHET_CHECK(launch_pass(true, list + list_base, list_len, list_edges, round));
                relaxed += static_cast<int64_t>(hm->v[3]);
                far_count = static_cast<int32_t>(hm->v[5 + sel]);
            }
        }
        settled_out += list_edges;
        if (far_count == 0) break;
        ++phases;
        // advance the threshold and split the far list
        double old_thr = threshold;
        {
            LevelInfo li;
            li.n = n; li.m = g->out.m; li.m_in = g->out.m; li.level = phases; li.frontier = list_len;
            li.frontier_edges = list_edges; li.unvisited = n - list_count; li.delta = base_delta;
            li.threshold = threshold; li.far_size = far_count;
            const int cls = decide(li);
            delta = base_delta * std::ldexp(1.0, cls - 1);
        }
        double new_thr = std::max(threshold + delta, std::nextafter(threshold, HUGE_VAL));
        list_base = list_count;
        for (;;) {
            const int other = 1 - sel;
            hipLaunchKernelGGL(k_ctr_reset, dim3(1), dim3(512), 0, s, ctr, other);
            const int32_t next_round = round + 1;
            unsigned grid = static_cast<unsigned>(
                std::max<int64_t>(1, std::min<int64_t>(ceil_div(far_count, kBT * kSplitItems), 1024)));
            const unsigned long long seq = mail_next(g);
            hipLaunchKernelGGL(k_far_split, dim3(grid), dim3(kBT), 0, s, far[sel], far_count, dist, old_thr, new_thr,
                               stamp_near, next_round, stamp_far, far_gen + 1, list, far[other], other, g->out.rowptr,
                               ctr, g->mail_dev, seq);
            GHET_HIP(hipGetLastError());
            GHET_CHECK(mail_wait(g, s, seq));
            list_count = static_cast<int32_t>(hm->v[7]);
            F = list_count - list_base;
            E = static_cast<int64_t>(hm->v[0]);
            far_count = static_cast<int32_t>(hm->v[5 + other]);
            const unsigned long long mb = hm->v[1];
            double far_min;
            std::memcpy(&far_min, &mb, sizeof(double));
            sel = other;
            ++far_gen;
            round = next_round;
            r0 = next_round;
            if (F > 0 || far_count == 0) break;
            // nothing in the next bucket: jump to the smallest far distance
            // (strictly above it, so the split always makes progress)
            old_thr = new_thr;
            new_thr = std::max(far_min + delta, std::nextafter(far_min, HUGE_VAL));
        }
        threshold = new_thr;
        list_len = F;
        list_edges = E;
        cur = list + list_base;
        if (F == 0 && far_count == 0) break;
    }
    if (pre_out) {
        GHET_CHECK(g->ws.get(WS_PRE, 4 * size_t(n), &p));
        int32_t* pre = static_cast<int32_t*>(p);
        GHET_CHECK(fill<int32_t>(pre, kIntInf, n, s));
        const unsigned gw = static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 4096)));
        hipLaunchKernelGGL(k_sssp_parent, dim3(gw), dim3(256), 0, s, n, g->out.rowptr, g->out.col, g->out.w, dist, inf,
                           src, pre);
        hipLaunchKernelGGL(k_pre_finish, dim3(gw), dim3(256), 0, s, pre, n);
    }
    GHET_HIP(hipEventRecord(g->ev_end, s));
    GHET_HIP(hipGetLastError());
    if (dist_out) GHET_CHECK(copy_out(g, dist_out, dist, 8 * size_t(n), s));
    if (pre_out) GHET_CHECK(copy_out(g, pre_out, g->ws.ptr[WS_PRE], 4 * size_t(n), s));
    if (st) {
        GHET_CHECK(reached_stats_f64(dist, inf, g->out.rowptr, n, ctr->slots, s));
        GHET_HIP(hipMemcpyAsync(g->pinned, ctr->slots, sizeof(CtrSlot) * kSlots, hipMemcpyDeviceToHost, s));
    }
    GHET_HIP(hipStreamSynchronize(s));
    if (st) {
        float ms = 0;
        GHET_HIP(hipEventElapsedTime(&ms, g->ev_begin, g->ev_end));
        st->device_ms = ms;
        st->levels = static_cast<int32_t>(rounds);
        const CtrSlot* hs = reinterpret_cast<const CtrSlot*>(g->pinned);
        st->vertices_reached = static_cast<int64_t>(slot_sum(hs, 0));
        st->edges_traversed = static_cast<int64_t>(slot_sum(hs, 1));
        st->trace_len = trace_len;
        st->extra[0] = base_delta;
        st->extra[1] = phases;
        st->extra[2] = static_cast<double>(relaxed);  // edges scanned over all passes
        st->extra[3] = pulls;                         // heavy passes run as pulls
        st->status_flags = sched.fallback ? 1 : 0;
    }
    return sched.fallback ? GHET_WARN_FALLBACK : GHET_OK;
}

}  // namespace ghet
