// Batched SSSP: up to 32 sources per pass advanced together in ONE persistent
// cooperative launch per pass (near-far buckets with a threshold shared by
// the batch, light/heavy edge split as in the single-source form).  The whole
// bucket loop runs on the device: rounds, heavy passes and far-list splits
// are separated by grid barriers, and the host only seeds the pass and
// collects its distances.
//
// No reference counterpart: the reference runs one csr_sssp per source
// (GPU_SSSP_csr.cuh:253), with host-driven rounds (:287-307).  Every source
// converges to the same fixpoint as its own run (the minimum over walks of
// the left-to-right rounded path sum, as for a single source), so results
// are bit-identical to per-source runs whatever the interleaving.
//
// Vertex queues with source masks.  A round's frontier is a queue of DISTINCT
// vertices; vertex u carries a 32-bit mask of the sources whose distance at
// u improved below the threshold.  One scan over the queue
// balances the round's edges for all sources at once, and each edge
// (u -> v, w) is read once and relaxed for every source in u's mask.
// Distances are vertex-major [n][K] (the K values of a vertex share cache
// lines).  Rows are ordered by weight (the single-source layout), so a round
// reads only the prefix of u's row that is light for the smallest masked
// distance, and relaxes per source only the edges with d_s[u] + w < thr; an
// improvement ORs the source's bit into v's next-round mask, whose
// 0 -> nonzero transition appends v to the next queue.  It also ORs the bit
// into v's bucket mask, whose 0 -> nonzero transition appends v to the
// bucket list.  When the near queue drains, one heavy pass over the bucket
// list relaxes, per bucket-mask source, the edges with d_s[u] + w >= thr
// (now final) and sends improvements to the far list through far masks.  The
// far list is then split at the next threshold per (vertex, source) bit; bits
// of entries already below the old threshold are stale and dropped.
//
// The caller's [k][n] rows are produced by one tiled transpose per pass.
#include <cmath>
#include <cstddef>
#include <cstring>

#include "graph_internal.hpp"

namespace ghet {

int sssp_base_delta(ghet_graph* g, double* out, bool size_term);
int build_sssp_layout(ghet_graph* g);

namespace {

// One phase's counters, in three sets: phase P counts in set P % 3 and clears
// set (P + 1) % 3, whose values every block read at the start of phase P - 1
// (before that phase's grid barriers).  Queue lengths are carried by every
// block in registers; a phase only counts its own appends.
struct alignas(64) PhaseCtr {
    int32_t near;              // vertices appended to the next near queue
    int32_t list;              // vertices appended to the bucket list
    int32_t far;               // vertices appended to a far list
    int32_t stop;              // set by block 0: phase or time budget spent
    int32_t tiles;             // relax tiles handed out beyond the first gridDim.x
    int32_t pad32;
    unsigned long long pad[5];
    CtrSlot slots[kSlots];     // v[1] smallest kept far distance (bits), v[3] edges scanned
};
static_assert(offsetof(PhaseCtr, slots) == 64, "phase_reset assumes 8 header words before the slots");
constexpr unsigned long long kInfBits = 0x7FF0000000000000ULL;

constexpr int kBT = 256;
constexpr int kItems = 8;
constexpr int kTile = kBT * kItems;
// LDS capacity of each staged queue (a tile pushes at most kTile entries; a
// flush empties it above kOut - kTile): 1.25 tiles instead of 2 lets 3
// blocks share a CU instead of 2.
constexpr int kOut = kTile + kTile / 4;
// a relax tile pushes at most kTile entries per queue, a split step kBT
static_assert(kOut >= kTile && kOut >= kBT, "staged queues smaller than one step's pushes");
constexpr int kMaxK = 32;                        // sources per pass (mask width)
constexpr int kMaxGrid = 4096;                   // blocks of the cooperative launch (btot entries)

__device__ __forceinline__ int upper_bound_i32(const int32_t* a, int len, int32_t key) {
    int lo = 0, hi = len;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (a[mid] <= key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

struct MssArgs {
    int32_t n, ks;                     // vertices, sources in this pass
    const int32_t* rowptr;             // out-CSR offsets
    const int32_t* col;                // weight-ordered rows (targets)
    const double* w;                   // weight-ordered rows (weights)
    double* dist;                      // [n][ks]
    uint32_t *mask_a, *mask_b;         // near masks of the current / next round
    uint32_t *fmask, *bmask;           // far masks, bucket masks
    int32_t *qa, *qb, *far0, *far1, *list;
    int32_t *scan, *start, *vert;      // a pass's entries: first edge, first position, vertex
    int32_t* tile_first;               // [ceil(m / kTile) + 2] entry holding each tile's first edge
    uint32_t* emask;                   // and source mask
    int32_t* btot;                     // [gridDim.x] edges of each block's entry range
    unsigned long long* bar;           // grid-barrier counters (kBarWords, zeroed before the launch)
    PhaseCtr* ctr;                     // [3]
    const int32_t* seed_v;             // distinct source vertices of the pass
    const uint32_t* seed_m;            // and their source masks
    int32_t seeds;
    double inf, delta;
    long long max_phases;              // budgets: the kernel stops (and reports it) past either
    unsigned long long max_cycles;     // wall-clock ticks
    long long* stats;                  // [0] rounds, [1] edges scanned, [2] stopped early
};

struct MssShared {
    int32_t scan[kTile], start[kTile], src[kTile];
    uint32_t mask[kTile];
    int32_t qa[kOut], qb[kOut];        // staged appends (relax: near, list / far; split: list, far)
    int32_t i0, span, na, nb, base;
    int64_t next_tile;
    int32_t wave[kBT / 64];
    unsigned long long red[kBT / 64];
    long long ctl[6];
};

// Block 0 clears a counter set (v[1] of every slot to +inf bits).
__device__ __forceinline__ void phase_reset(PhaseCtr* c) {
    unsigned long long* w = reinterpret_cast<unsigned long long*>(c);
    constexpr int kWords = sizeof(PhaseCtr) / 8;
    for (int i = threadIdx.x; i < kWords; i += kBT) w[i] = (i >= 8 && (i - 8) % 8 == 1) ? kInfBits : 0ULL;
}

struct PhaseRes {
    int32_t near, list, far, stop;
    unsigned long long dmin, scanned;
};

// The finished phase's counts (block-uniform; call after a grid barrier).
__device__ __forceinline__ PhaseRes phase_read(const PhaseCtr* c, MssShared& sh) {
    if (threadIdx.x < 64) {  // wave 0: lane i reads slot i
        const CtrSlot* sl = c->slots + threadIdx.x;
        const unsigned long long mn = wave_min(load_agent(&sl->v[1]));
        const unsigned long long sc = wave_sum(load_agent(&sl->v[3]));
        if (threadIdx.x == 0) {
            auto ld = [](const int32_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); };
            sh.ctl[0] = ld(&c->near);
            sh.ctl[1] = ld(&c->list);
            sh.ctl[2] = ld(&c->far);
            sh.ctl[3] = ld(&c->stop);
            sh.ctl[4] = static_cast<long long>(mn);
            sh.ctl[5] = static_cast<long long>(sc);
        }
    }
    __syncthreads();
    PhaseRes r;
    r.near = static_cast<int32_t>(sh.ctl[0]);
    r.list = static_cast<int32_t>(sh.ctl[1]);
    r.far = static_cast<int32_t>(sh.ctl[2]);
    r.stop = static_cast<int32_t>(sh.ctl[3]);
    r.dmin = static_cast<unsigned long long>(sh.ctl[4]);
    r.scanned = static_cast<unsigned long long>(sh.ctl[5]);
    __syncthreads();
    return r;
}

// Row segments of the F queue entries q[] (masks qm[], consumed and cleared
// here): round, the prefix with min_s d_s[u] + w < thr, which holds every
// masked source's light edges; heavy, the rest from the prefix end of
// max_s d_s[u], which holds every masked source's heavy edges.  Entry i's
// edge count goes to scan[i], its first position to start[i], its vertex and
// mask to vert[i] and emask[i].  Block b takes the contiguous entries
// [F b / G, F (b + 1) / G), one per thread at a time, and writes their total
// to btot[b]; mss_scan turns the counts into offsets after a grid barrier.
// Queues of at most kBT entries are done by block 0 alone, scan included
// (returns true there: no barrier and no mss_scan needed).
__device__ __forceinline__ int64_t range_lo(int32_t F, int64_t b) { return static_cast<int64_t>(F) * b / gridDim.x; }

template <bool kHeavy>
__device__ __forceinline__ int32_t row_segment(const MssArgs& a, int32_t u, uint32_t m, double thr, int32_t* st) {
    const double* du = a.dist + static_cast<int64_t>(u) * a.ks;
    double d = kHeavy ? 0.0 : HUGE_VAL;
    for (uint32_t rem = m; rem;) {  // four loads in flight at a time
        double x[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int sx = __ffs(rem) - 1;
            rem &= rem - 1;
            x[j] = sx >= 0 ? du[sx] : (kHeavy ? 0.0 : HUGE_VAL);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) d = kHeavy ? fmax(d, x[j]) : fmin(d, x[j]);
    }
    const int32_t rb = a.rowptr[u], re = a.rowptr[u + 1];
    int32_t lo = rb, hi = re;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (d + a.w[mid] < thr) lo = mid + 1;
        else hi = mid;
    }
    *st = kHeavy ? lo : rb;
    return kHeavy ? re - lo : lo - rb;
}

// Tile index of a pass: tile_first[k] = the entry holding edge k * kTile (an
// entry with edges [o, o + c) marks every k with o <= k kTile < o + c), and
// tile_first[ceil(E / kTile)] = F - 1 closes the last tile; a relax tile's
// entries are then tile_first[t] .. tile_first[t + 1], with no binary search
// of the scan.
__device__ __forceinline__ void mark_tiles(int32_t* tile_first, int32_t slot, int32_t o, int32_t c) {
    for (int64_t k = (static_cast<int64_t>(o) + kTile - 1) / kTile; k * kTile < static_cast<int64_t>(o) + c; ++k)
        tile_first[k] = slot;
}

// Exclusive scan of one value per thread over the block; *total gets the sum.
__device__ __forceinline__ int32_t block_excl_scan(int32_t x, int32_t* total, MssShared& sh) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int32_t incl = x;
    for (int off = 1; off < 64; off <<= 1) {
        const int32_t y = __shfl_up(incl, off);
        if (lane >= off) incl += y;
    }
    if (lane == 63) sh.wave[wave] = incl;
    __syncthreads();
    int32_t run = incl - x, t = 0;
    for (int v = 0; v < kBT / 64; ++v) {
        if (v < wave) run += sh.wave[v];
        t += sh.wave[v];
    }
    __syncthreads();  // sh.wave is reused by the next call
    *total = t;
    return run;
}

template <bool kHeavy>
__device__ bool mss_count(const MssArgs& a, const int32_t* __restrict__ q, int32_t F, uint32_t* qm, double thr,
                          MssShared& sh) {
    const int tid = threadIdx.x;
    if (F <= kBT) {
        if (blockIdx.x == 0) {
            int32_t cnt = 0, st = 0, u = 0;
            uint32_t m = 0;
            if (tid < F) {
                u = q[tid];
                m = qm[u];
                qm[u] = 0u;
                cnt = row_segment<kHeavy>(a, u, m, thr, &st);
            }
            int32_t total;
            const int32_t off = block_excl_scan(cnt, &total, sh);
            if (tid < F) {
                a.scan[tid] = off;
                a.start[tid] = st;
                a.vert[tid] = u;
                a.emask[tid] = m;
                mark_tiles(a.tile_first, tid, off, cnt);
            }
            if (tid == 0) {
                a.scan[F] = total;
                a.tile_first[(static_cast<int64_t>(total) + kTile - 1) / kTile] = F - 1;
            }
        }
        return true;
    }
    const int64_t b0 = range_lo(F, blockIdx.x), b1 = range_lo(F, blockIdx.x + 1);
    int32_t sum = 0;
    for (int64_t i = b0 + tid; i < b1; i += kBT) {
        const int32_t u = q[i];
        const uint32_t m = qm[u];
        qm[u] = 0u;
        int32_t st;
        const int32_t cnt = row_segment<kHeavy>(a, u, m, thr, &st);
        a.scan[i] = cnt;
        a.start[i] = st;
        a.vert[i] = u;
        a.emask[i] = m;
        sum += cnt;
    }
    int32_t total;
    (void)block_excl_scan(sum, &total, sh);
    if (tid == 0) a.btot[blockIdx.x] = total;
    return false;
}

// After a grid barrier: block b adds the totals of blocks before it and
// scans its range's counts in place; the last block writes scan[F].
__device__ void mss_scan(const MssArgs& a, int32_t F, MssShared& sh) {
    const int tid = threadIdx.x;
    int32_t part = 0;
    for (int j = tid; j < static_cast<int>(blockIdx.x); j += kBT) part += a.btot[j];
    int32_t carry;
    (void)block_excl_scan(part, &carry, sh);
    const int64_t b0 = range_lo(F, blockIdx.x), b1 = range_lo(F, blockIdx.x + 1);
    for (int64_t base = b0; base < b1; base += kBT) {
        const int64_t i = base + tid;
        const int32_t c = i < b1 ? a.scan[i] : 0;
        int32_t t;
        const int32_t off = block_excl_scan(c, &t, sh);
        if (i < b1) {
            a.scan[i] = carry + off;
            mark_tiles(a.tile_first, static_cast<int32_t>(i), carry + off, c);
        }
        carry += t;
    }
    if (blockIdx.x == gridDim.x - 1 && tid == 0) {
        a.scan[F] = carry;
        a.tile_first[(static_cast<int64_t>(carry) + kTile - 1) / kTile] = F - 1;
    }
}

// One pass over the prepared entries, in kTile-edge tiles; entry j covers
// edges [start[j], start[j] + scan[j+1] - scan[j]) of the weight-ordered
// rows.  Round (kHeavy false): per source in the entry's mask, relaxes the
// edges with d_s[u] + w < thr; an improvement ORs the source's bit into v's
// next-round mask (0 -> nonzero appends v to next_q) and its bucket mask
// (0 -> nonzero appends v to list_out).  Heavy: relaxes the edges with
// d_s[u] + w >= thr; improvements go to v's far mask (0 -> nonzero appends v
// to far_out).
template <bool kHeavy>
__device__ void mss_relax(const MssArgs& a, int32_t F, double thr, uint32_t* mask_next, int32_t* next_q,
                          int32_t* list_out, int32_t* far_out, PhaseCtr* c, MssShared& sh) {
    const int tid = threadIdx.x;
    const int32_t E = __hip_atomic_load(a.scan + F, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (E == 0) return;  // block-uniform
    if (tid == 0) {
        sh.na = 0;
        sh.nb = 0;
    }
    __syncthreads();
    auto nop = [](int32_t) {};
    if (blockIdx.x == 0 && tid == 0) slot_add(c->slots, 3, static_cast<unsigned long long>(E));
    // Tiles are handed out dynamically: block b starts with tile b, then
    // takes the next unclaimed one (one atomic per tile), so blocks that
    // draw cheap tiles take more of them.
    const int64_t ntiles = ceil_div(E, kTile);
    for (int64_t tile = blockIdx.x; tile < ntiles;) {
        const int32_t t0 = static_cast<int32_t>(tile * kTile);
        const int32_t t1 = static_cast<int32_t>(std::min<int64_t>(static_cast<int64_t>(t0) + kTile, E));
        if (tid == 0) {
            const int i0 = a.tile_first[tile];      // holds edge t0
            const int i1 = a.tile_first[tile + 1];  // holds edge t1, or the pass's last entry
            sh.i0 = i0;
            sh.span = i1 - i0 + 1;
            sh.next_tile = static_cast<int64_t>(gridDim.x) + atomicAdd(&c->tiles, 1);
        }
        __syncthreads();
        const int i0 = sh.i0, span = sh.span;
        const int64_t next_tile = sh.next_tile;
        const bool use_lds = span <= kTile;
        if (use_lds)
            for (int j = tid; j < span; j += kBT) {
                sh.scan[j] = a.scan[i0 + j];
                sh.start[j] = a.start[i0 + j];
                sh.src[j] = a.vert[i0 + j];
                sh.mask[j] = a.emask[i0 + j];
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
                    const int j = upper_bound_i32(sh.scan, span, e) - 1;
                    pos[k] = sh.start[j] + (e - sh.scan[j]);
                    u[k] = sh.src[j];
                    msk[k] = sh.mask[j];
                } else {
                    const int j = upper_bound_i32(a.scan + i0, span, e) - 1;
                    u[k] = a.vert[i0 + j];
                    pos[k] = a.start[i0 + j] + (e - a.scan[i0 + j]);
                    msk[k] = a.emask[i0 + j];
                }
            }
        }
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            v[k] = a.col[pos[k]];
            w[k] = a.w[pos[k]];
        }
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            bool push_a = false, push_b = false;
            const double* du = a.dist + static_cast<int64_t>(u[k]) * a.ks;
            double* dv = a.dist + static_cast<int64_t>(v[k]) * a.ks;
            // the masked sources four at a time: their distance loads, then
            // their atomics, are in flight together
            uint32_t bits = 0;
            for (uint32_t rem = msk[k]; rem;) {
                int sx[4];
                double nd[4], cur[4], old[4];
                bool c[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    sx[j] = __ffs(rem) - 1;  // -1 once the mask is spent
                    rem &= rem - 1;
                }
#pragma unroll
                for (int j = 0; j < 4; ++j) nd[j] = sx[j] >= 0 ? du[sx[j]] + w[k] : 0.0;
#pragma unroll
                for (int j = 0; j < 4; ++j) c[j] = sx[j] >= 0 && (kHeavy ? nd[j] >= thr : nd[j] < thr);
#pragma unroll
                for (int j = 0; j < 4; ++j) cur[j] = c[j] ? dv[sx[j]] : 0.0;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    c[j] = c[j] && nd[j] < cur[j];
                    old[j] = c[j] ? atomic_min_nonneg(&dv[sx[j]], nd[j]) : 0.0;
                }
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (c[j] && nd[j] < old[j]) bits |= 1u << sx[j];
            }
            if (bits) {
                if (kHeavy) {
                    push_a = atomicOr(&a.fmask[v[k]], bits) == 0u;
                } else {
                    push_a = atomicOr(&mask_next[v[k]], bits) == 0u;
                    push_b = atomicOr(&a.bmask[v[k]], bits) == 0u;
                }
            }
            lq_push(sh.qa, &sh.na, push_a, v[k]);
            if (!kHeavy) lq_push(sh.qb, &sh.nb, push_b, v[k]);
        }
        if (kHeavy) {
            lq_flush<kBT>(sh.qa, &sh.na, &sh.base, &c->far, far_out, kOut - kTile, nop);
        } else {
            lq_flush<kBT>(sh.qa, &sh.na, &sh.base, &c->near, next_q, kOut - kTile, nop);
            lq_flush<kBT>(sh.qb, &sh.nb, &sh.base, &c->list, list_out, kOut - kTile, nop);
        }
        tile = next_tile;  // lq_flush's barriers order this read before sh.next_tile's next write
    }
    if (kHeavy) {
        lq_flush<kBT>(sh.qa, &sh.na, &sh.base, &c->far, far_out, 0, nop);
    } else {
        lq_flush<kBT>(sh.qa, &sh.na, &sh.base, &c->near, next_q, 0, nop);
        lq_flush<kBT>(sh.qb, &sh.nb, &sh.base, &c->list, list_out, 0, nop);
    }
}

// Splits the far list at new_thr: per (vertex, source) bit, entries below
// old_thr are stale (settled in an earlier bucket) and dropped, entries below
// new_thr start the next bucket, the rest stay far.  A vertex with bits in
// the new bucket is appended to the bucket list, which is also the bucket's
// first near queue (its near and bucket masks get those bits; both masks are
// all-zero here and far-list entries are distinct).  The smallest kept far
// distance goes to the slots' v[1].
__device__ void mss_split(const MssArgs& a, const int32_t* __restrict__ far_in, int32_t count, double old_thr,
                          double new_thr, uint32_t* mask_near, int32_t* far_out, PhaseCtr* c, MssShared& sh) {
    if (threadIdx.x == 0) {
        sh.na = 0;
        sh.nb = 0;
    }
    __syncthreads();
    auto nop = [](int32_t) {};
    unsigned long long dmin = kInfBits;
    for (int64_t base = static_cast<int64_t>(blockIdx.x) * kBT; base < count;
         base += static_cast<int64_t>(gridDim.x) * kBT) {
        const int64_t i = base + threadIdx.x;
        bool to_near = false, keep = false;
        int32_t v = 0;
        if (i < count) {
            v = far_in[i];
            const uint32_t m = a.fmask[v];
            const double* dv = a.dist + static_cast<int64_t>(v) * a.ks;
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
            a.fmask[v] = kb;
            keep = kb != 0u;
            if (nb) {
                mask_near[v] = nb;
                a.bmask[v] = nb;
                to_near = true;
            }
        }
        lq_push(sh.qa, &sh.na, to_near, v);
        lq_push(sh.qb, &sh.nb, keep, v);
        lq_flush<kBT>(sh.qa, &sh.na, &sh.base, &c->list, a.list, kOut - kBT, nop);
        lq_flush<kBT>(sh.qb, &sh.nb, &sh.base, &c->far, far_out, kOut - kBT, nop);
    }
    lq_flush<kBT>(sh.qa, &sh.na, &sh.base, &c->list, a.list, 0, nop);
    lq_flush<kBT>(sh.qb, &sh.nb, &sh.base, &c->far, far_out, 0, nop);
    dmin = wave_min(dmin);
    if ((threadIdx.x & 63) == 0) sh.red[threadIdx.x >> 6] = dmin;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 0; w < kBT / 64; ++w) dmin = sh.red[w] < dmin ? sh.red[w] : dmin;
        if (dmin != kInfBits) atomicMin(&c->slots[blockIdx.x & (kSlots - 1)].v[1], dmin);
    }
    __syncthreads();
}

// One pass of up to 32 sources in one cooperative launch: the whole bucket
// loop (light rounds, the heavy pass, far-list splits) runs on the device,
// with grid barriers between the steps.  Every block reads the same counts
// after each barrier, so all take the same branches and meet at every
// barrier.  A phase or wall-clock budget, checked by block 0 and published
// through the phase's counters, ends the loop uniformly (stats[2] = 1).
__global__ void __launch_bounds__(kBT) k_mss_run(MssArgs a) {
    __shared__ MssShared sh;
    unsigned long long epoch = 0;  // grid barriers passed
    const int tid = threadIdx.x;
    const int64_t gthreads = static_cast<int64_t>(gridDim.x) * kBT;
    const int64_t gtid = static_cast<int64_t>(blockIdx.x) * kBT + tid;
    const int64_t kn = static_cast<int64_t>(a.n) * a.ks;
    for (int64_t i = gtid; i < kn; i += gthreads) a.dist[i] = a.inf;
    for (int64_t i = gtid; i < a.n; i += gthreads) {
        a.mask_a[i] = 0u;
        a.mask_b[i] = 0u;
        a.fmask[i] = 0u;
        a.bmask[i] = 0u;
    }
    if (blockIdx.x == 0) {
        phase_reset(a.ctr);
        phase_reset(a.ctr + 1);
    }
    grid_barrier(a.bar, epoch);
    if (blockIdx.x == 0)
        for (int j = tid; j < a.seeds; j += kBT) {
            const int32_t v = a.seed_v[j];
            const uint32_t m = a.seed_m[j];
            a.mask_a[v] = m;
            a.bmask[v] = m;
            a.list[j] = v;
            for (uint32_t b = m; b; b &= b - 1) a.dist[static_cast<int64_t>(v) * a.ks + (__ffs(b) - 1)] = 0.0;
        }
    grid_barrier(a.bar, epoch);

    uint32_t *mcur = a.mask_a, *mnext = a.mask_b;
    const int32_t* cur = a.list;  // a bucket's first near queue is its list
    int32_t* nxt = a.qa;
    int32_t* far[2] = {a.far0, a.far1};
    int32_t F = a.seeds, list_count = a.seeds, far_count[2] = {0, 0};
    int sel = 0;
    double thr = a.delta;
    long long P = 0, rounds = 0;
    unsigned long long scanned = 0;
    bool stopped = false;
    const unsigned long long t0 = wall_clock64();
    // Start of a phase: block 0 clears the set after this phase's and flags
    // a spent budget in this phase's set.
    auto begin_phase = [&]() -> PhaseCtr* {
        PhaseCtr* c = a.ctr + (P % 3);
        if (blockIdx.x == 0) {
            phase_reset(a.ctr + ((P + 1) % 3));
            if (tid == 0 && (P + 1 >= a.max_phases || wall_clock64() - t0 > a.max_cycles)) c->stop = 1;
        }
        return c;
    };
    for (;;) {
        // light rounds of the current bucket
        while (F > 0) {
            PhaseCtr* c = begin_phase();
            if (!mss_count<false>(a, cur, F, mcur, thr, sh)) {
                grid_barrier(a.bar, epoch);
                mss_scan(a, F, sh);
            }
            grid_barrier(a.bar, epoch);
            mss_relax<false>(a, F, thr, mnext, nxt, a.list + list_count, nullptr, c, sh);
            grid_barrier(a.bar, epoch);
            const PhaseRes r = phase_read(c, sh);
            ++P;
            ++rounds;
            scanned += r.scanned;
            F = r.near;
            list_count += r.list;
            cur = nxt;
            nxt = (cur == a.qa) ? a.qb : a.qa;
            uint32_t* t = mcur;
            mcur = mnext;
            mnext = t;
            if (r.stop) {
                stopped = true;
                break;
            }
        }
        if (stopped) break;
        // heavy pass over the drained bucket: its distances are final
        if (list_count > 0) {
            PhaseCtr* c = begin_phase();
            if (!mss_count<true>(a, a.list, list_count, a.bmask, thr, sh)) {
                grid_barrier(a.bar, epoch);
                mss_scan(a, list_count, sh);
            }
            grid_barrier(a.bar, epoch);
            mss_relax<true>(a, list_count, thr, nullptr, nullptr, nullptr, far[sel] + far_count[sel], c, sh);
            grid_barrier(a.bar, epoch);
            const PhaseRes r = phase_read(c, sh);
            ++P;
            scanned += r.scanned;
            far_count[sel] += r.far;
            if (r.stop) {
                stopped = true;
                break;
            }
        }
        if (far_count[sel] == 0) break;
        // split the far list at the next threshold into the next bucket's
        // list (its near masks go to mcur, all-zero after the last round)
        double old_thr = thr;
        double new_thr = fmax(thr + a.delta, nextafter(thr, HUGE_VAL));
        for (;;) {
            const int other = 1 - sel;
            PhaseCtr* c = begin_phase();
            mss_split(a, far[sel], far_count[sel], old_thr, new_thr, mcur, far[other], c, sh);
            grid_barrier(a.bar, epoch);
            const PhaseRes r = phase_read(c, sh);
            ++P;
            ++rounds;
            F = r.list;
            list_count = r.list;
            far_count[other] = r.far;
            far_count[sel] = 0;
            sel = other;
            if (r.stop) {
                stopped = true;
                break;
            }
            if (F > 0 || far_count[sel] == 0) break;
            // nothing in the next bucket: jump to the smallest far distance
            // (strictly above it, so the split always makes progress)
            const double far_min = __longlong_as_double(static_cast<long long>(r.dmin));
            old_thr = new_thr;
            new_thr = fmax(far_min + a.delta, nextafter(far_min, HUGE_VAL));
        }
        if (stopped) break;
        thr = new_thr;
        cur = a.list;
        nxt = a.qa;
        if (F == 0 && far_count[sel] == 0) break;
    }
    if (blockIdx.x == 0 && tid == 0) {
        a.stats[0] = rounds;
        a.stats[1] = static_cast<long long>(scanned);
        a.stats[2] = stopped ? 1 : 0;
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
    GHET_CHECK(build_sssp_layout(g));
    const SsspLayout& L = g->sl;
    MssArgs a;
    a.n = n;
    a.rowptr = g->out.rowptr;
    a.col = L.col;
    a.w = L.w;
    a.inf = inf;
    a.delta = delta;
    GHET_CHECK(g->ws.get(WS_MULTI_A, 8 * size_t(KN), &p)); a.dist = static_cast<double*>(p);
    GHET_CHECK(g->ws.get(WS_MULTI_B, 16 * size_t(n), &p));
    a.mask_a = static_cast<uint32_t*>(p);
    a.mask_b = a.mask_a + n;
    a.fmask = a.mask_b + n;
    a.bmask = a.fmask + n;
    GHET_CHECK(g->ws.get(WS_MULTI_C, 32 * size_t(n), &p));
    a.qa = static_cast<int32_t*>(p);
    a.qb = a.qa + n;
    a.far0 = a.qb + n;
    a.far1 = a.far0 + n;
    a.list = a.far1 + n;
    a.start = a.list + n;
    a.vert = a.start + n;
    a.emask = reinterpret_cast<uint32_t*>(a.vert + n);
    GHET_CHECK(g->ws.get(WS_MULTI_D, 4 * (size_t(n) + 1), &p)); a.scan = static_cast<int32_t*>(p);
    // a pass's entries are distinct vertices: at most m edges
    GHET_CHECK(g->ws.get(WS_SSSP_TILE, 4 * (size_t(ceil_div(g->out.m, kTile)) + 2), &p));
    a.tile_first = static_cast<int32_t*>(p);
    GHET_CHECK(g->ws.get(WS_COUNTERS, 3 * sizeof(PhaseCtr) + 64, &p));
    a.ctr = static_cast<PhaseCtr*>(p);
    a.stats = reinterpret_cast<long long*>(a.ctr + 3);
    GHET_CHECK(g->ws.get(WS_MISC, 8 * size_t(kMaxK) + 4 * size_t(kMaxGrid) + 256, &p));
    int32_t* d_q = static_cast<int32_t*>(p);
    uint32_t* d_qmask = reinterpret_cast<uint32_t*>(d_q + kMaxK);
    a.seed_v = d_q;
    a.seed_m = d_qmask;
    a.btot = d_q + 2 * kMaxK;
    GHET_CHECK(g->ws.get(WS_BAR, 8 * size_t(kBarWords), &p));
    a.bar = static_cast<unsigned long long*>(p);
    double* tr = nullptr;
    if (dist_out) {
        GHET_CHECK(g->ws.get(WS_PR_OUT, 8 * size_t(KN), &p));
        tr = static_cast<double*>(p);
    }
    // Grid: every block resident at once (the launch is cooperative).
    int occ = 0, cus = 0, clock_khz = 0;
    GHET_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(k_mss_run), kBT, 0));
    GHET_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device));
    GHET_HIP(hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, g->device));
    if (occ < 1) return set_error(GHET_ERR_HIP, "batched SSSP kernel cannot be resident (occupancy 0)");
    const unsigned grid = static_cast<unsigned>(std::min(occ * cus, kMaxGrid));
    // Budgets that end a pass that does not converge (a logic error, never a
    // valid input): phases, and two minutes of device time.  reserved[5]
    // overrides the phase budget (tests force a small one).
    a.max_phases = (have_opts && o->reserved[5] > 0) ? o->reserved[5] : 64 * (static_cast<long long>(n) + 1024);
    a.max_cycles = static_cast<unsigned long long>(std::max(clock_khz, 1)) * 1000ULL * 120ULL;
    const long long* hstats = reinterpret_cast<const long long*>(g->pinned);
    const CtrSlot* hslots = reinterpret_cast<const CtrSlot*>(reinterpret_cast<const char*>(g->pinned) + 64);
    static_assert(64 + sizeof(CtrSlot) * kSlots <= kPinnedBytes, "pinned buffer");

    int64_t traversed = 0, rounds = 0, scanned = 0;
    GHET_HIP(hipEventRecord(g->ev_begin, s));
    for (int32_t p0 = 0; p0 < k; p0 += K) {
        const int32_t kp = std::min(K, k - p0);
        const int64_t kn = int64_t(kp) * n;
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
        // pageable sources: hipMemcpyAsync stages them before returning
        GHET_HIP(hipMemcpyAsync(d_q, hq.data(), 4 * hq.size(), hipMemcpyHostToDevice, s));
        GHET_HIP(hipMemcpyAsync(d_qmask, hm.data(), 4 * hm.size(), hipMemcpyHostToDevice, s));
        a.ks = kp;
        a.seeds = static_cast<int32_t>(hq.size());
        GHET_CHECK(fill<uint64_t>(reinterpret_cast<uint64_t*>(a.bar), 0, kBarWords, s));
        void* args[] = {&a};
        GHET_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_mss_run), dim3(grid), dim3(kBT), args, 0, s));
        GHET_HIP(hipMemcpyAsync(g->pinned, a.stats, 3 * sizeof(long long), hipMemcpyDeviceToHost, s));
        GHET_HIP(hipStreamSynchronize(s));
        if (hstats[2])
            return set_error(GHET_ERR_HIP, "batched SSSP: pass at source %d stopped after %lld rounds (phase or time "
                             "budget spent); results are invalid", p0, static_cast<long long>(hstats[0]));
        rounds += hstats[0];
        scanned += hstats[1];
        if (st) {
            GHET_CHECK(fill<uint64_t>(reinterpret_cast<uint64_t*>(a.ctr->slots), 0, sizeof(CtrSlot) / 8 * kSlots, s));
            hipLaunchKernelGGL(k_mss_reached, dim3(2048), dim3(256), 0, s, a.dist, kn, kp, inf, g->out.rowptr,
                               a.ctr->slots);
            GHET_HIP(hipMemcpyAsync(reinterpret_cast<char*>(g->pinned) + 64, a.ctr->slots, sizeof(CtrSlot) * kSlots,
                                    hipMemcpyDeviceToHost, s));
        }
        if (dist_out) {
            hipLaunchKernelGGL(k_mss_transpose,
                               dim3(static_cast<unsigned>(ceil_div(n, 64)), static_cast<unsigned>(ceil_div(kp, 64))),
                               dim3(64, 4), 0, s, a.dist, n, kp, tr);
            GHET_HIP(hipGetLastError());
            GHET_HIP(hipMemcpyAsync(dist_out + int64_t(p0) * n, tr, 8 * size_t(kn), hipMemcpyDeviceToHost, s));
        }
        GHET_HIP(hipStreamSynchronize(s));
        if (st) traversed += static_cast<int64_t>(slot_sum(hslots, 2));
    }
    GHET_HIP(hipEventRecord(g->ev_end, s));
    GHET_HIP(hipStreamSynchronize(s));
    if (st) {
        float ms = 0;
        GHET_HIP(hipEventElapsedTime(&ms, g->ev_begin, g->ev_end));
        st->device_ms = ms;
        st->levels = static_cast<int32_t>(rounds);
        st->hot_kernel_launches = static_cast<int32_t>(ceil_div(k, K));  // one cooperative launch per pass
        st->edges_traversed = traversed;
        st->extra[0] = delta;
        st->extra[1] = K;
        st->extra[2] = static_cast<double>(scanned);  // edges scanned over all passes
    }
    return GHET_OK;
}

}  // namespace ghet
