// Community detection by label propagation (CDLP) on gfx950.
//
// Reference: CDLP_GPU (GPU_CDLP_csr.cuh:75-187) copies every neighbour label
// into a 2m-entry array, segment-sorts it with cub::DeviceSegmentedSort and
// takes the longest run; CDLP (CPU_Community_Detection.hpp:12-64) counts
// labels of in- and out-neighbours (with multiplicity) in a hash map and takes
// the most frequent, ties to the smallest label; a vertex without neighbours
// keeps its label.  (The reference GPU kernel reads the next vertex's segment
// for an isolated vertex, GPU_CDLP_csr.cuh:56; this path follows the CPU and
// LDBC semantics.)
//
// No sort: vertices are bucketed by all-degree d once per graph and each class
// counts labels where it fits best:
//   d <= 32      a wave gathers the labels of 64 such vertices into LDS
//                together, then each edge's lane counts its label over its
//                vertex's labels and offers the result to the vertex's slot;
//   d <= 256     one wave, labels in LDS, counted in the wave's 512-slot
//                open-addressing hash table;
//   d <= 8192    one 256-thread block, open-addressing hash table in LDS
//                (the power of two >= 2d slots of a 4096-slot table; above
//                2048 neighbours, sub-passes if the labels overflow it);
//   d > 8192     hubs: every 8192-edge chunk counts its labels in an LDS
//                table and writes the distinct (label, count) pairs to its
//                own region; a plan gives each hub ceil(pairs / 4096) merge
//                items, item (hub, p) adds up the pairs whose label hashes to
//                partition p in an LDS table and offers its winner with one
//                64-bit atomicMax.  No global hash table, no table scan.
// Winner = max count, ties to the smallest label: the maximum of the key
// (count << 32) | (0xFFFFFFFF - label).  The first iteration, from identity
// labels, has a closed form and skips the counting (k_cdlp_first_*).
#include "graph_internal.hpp"

#include <rocprim/rocprim.hpp>

namespace ghet {
namespace {

constexpr int kBT = 256;
constexpr int kSmall = 32, kWave = 256, kLds = 2048, kLdsSlots = 4096;
constexpr int kBig = 8192, kBigSlots = 16384, kBigBT = 1024;  // class 3 and the hub kernels
constexpr int kChunk = 8192;          // edges per hub chunk (at most 8192 distinct labels: half a table)
constexpr int kMergeTarget = 4096;    // pairs per hub merge item
constexpr int kMaxHubChunks = 4096;   // chunks of one hub the merge kernel's LDS prefix holds
constexpr int32_t kEmpty = -1;
constexpr int kLdsLoads = 4;                // edges per lane loaded together in k_cdlp_lds
constexpr int kHubLoads = kChunk / kBigBT;  // a hub chunk's edges per lane, loaded together
static_assert(kHubLoads * kBigBT == kChunk, "a hub chunk is whole steps of the block");

__device__ __forceinline__ uint64_t pack(uint32_t count, int32_t label) {
    return (static_cast<uint64_t>(count) << 32) | (0xFFFFFFFFu - static_cast<uint32_t>(label));
}
__device__ __forceinline__ int32_t unpack_label(uint64_t k) {
    return static_cast<int32_t>(0xFFFFFFFFu - static_cast<uint32_t>(k & 0xFFFFFFFFu));
}
__device__ __forceinline__ uint32_t hash_slot(int32_t label, uint32_t mask) {
    return (static_cast<uint32_t>(label) * 0x9E3779B1u) & mask;
}

__global__ void k_label_init(int32_t* a, int32_t* b, int32_t n) {
    for (int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < n;
         v += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        a[v] = static_cast<int32_t>(v);
        b[v] = static_cast<int32_t>(v);
    }
}

// Per-wave tiles for the lane class.  A wave takes a tile of 64 consecutive
// list entries (at most 2048 labels), scans their degrees and gathers all
// their labels into its LDS buffer in edge order: lane l takes the tile's
// edges l, l + 64, ..., finds each edge's vertex by binary search over the
// tile's offsets and issues kTileLoads edges' loads together.  The counting
// then reads LDS only.
constexpr int kTileLabels = 2048;
constexpr int kTileLoads = 8;

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
}

// The tile's vertex of lane `lane` < nv is (b, d); lanes past the tile pass
// d = 0.  TILE (a power of two) entries of s_off/s_b are used.  Returns the
// lane's first label position; *total = the tile's labels.
template <int TILE>
__device__ __forceinline__ int32_t tile_gather(int32_t b, int32_t d, int32_t* s_off, int32_t* s_b, int32_t* s_lab,
                                               const int32_t* __restrict__ col, const int32_t* __restrict__ lab,
                                               int32_t* total) {
    const int lane = lane_id();
    int32_t incl = d;
    for (int o = 1; o < 64; o <<= 1) {
        const int32_t y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    const int32_t off = incl - d;
    const int32_t E = __shfl(incl, 63);
    if (lane < TILE) {
        s_off[lane] = off;
        s_b[lane] = b;
    }
    wave_lds_sync();
    for (int32_t k0 = 0; k0 < E; k0 += 64 * kTileLoads) {
        int32_t ce[kTileLoads];
#pragma unroll
        for (int u = 0; u < kTileLoads; ++u) {
            const int32_t k = k0 + u * 64 + lane;
            int j = 0;  // last tile vertex with s_off[j] <= k (degrees are >= 1)
#pragma unroll
            for (int step = TILE / 2; step >= 1; step >>= 1) j += s_off[j + step] <= k ? step : 0;
            ce[u] = k < E ? s_b[j] + (k - s_off[j]) : 0;  // lanes past the tile load edge 0
        }
        int32_t cv[kTileLoads];
#pragma unroll
        for (int u = 0; u < kTileLoads; ++u) cv[u] = col[ce[u]];
#pragma unroll
        for (int u = 0; u < kTileLoads; ++u) {
            const int32_t k = k0 + u * 64 + lane;
            const int32_t x = lab[cv[u]];
            if (k < E) s_lab[k] = x;
        }
    }
    wave_lds_sync();
    *total = E;
    return off;
}

// d <= 32: tiles of 64 list entries (a lane per vertex), labels gathered
// together, then counted edge-parallel.
__global__ void __launch_bounds__(kBT) k_cdlp_small(const int32_t* __restrict__ list, int32_t count,
                                                    const int32_t* __restrict__ rowptr,
                                                    const int32_t* __restrict__ col, const int32_t* __restrict__ lab,
                                                    int32_t* __restrict__ out) {
    __shared__ int32_t s_lab[kBT / 64][kTileLabels];
    __shared__ int32_t s_off[kBT / 64][64];
    __shared__ int32_t s_b[kBT / 64][64];
    __shared__ unsigned long long s_best[kBT / 64][64];
    const int lane = lane_id(), w = threadIdx.x >> 6;
    const int64_t tiles = ceil_div(count, 64);
    for (int64_t t = (static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x) >> 6; t < tiles;
         t += (static_cast<int64_t>(gridDim.x) * kBT) >> 6) {
        const int64_t i = t * 64 + lane;
        const int32_t v = i < count ? list[i] : 0;
        const int32_t b = i < count ? rowptr[v] : 0;
        const int32_t d = i < count ? rowptr[v + 1] - b : 0;
        int32_t E;
        s_best[w][lane] = 0ull;
        const int32_t off = tile_gather<64>(b, d, s_off[w], s_b[w], s_lab[w], col, lab, &E);
        // edge-parallel counting: the lane of edge k counts its label over
        // k's vertex's labels and offers (count, label) to that vertex's
        // slot; the work per lane is the sum of its edges' degrees, not the
        // square of the tile's largest degree
        for (int32_t k = lane; k < E; k += 64) {
            int j = 0;
#pragma unroll
            for (int step = 32; step >= 1; step >>= 1) j += s_off[w][j + step] <= k ? step : 0;
            const int32_t so = s_off[w][j];
            const int32_t se = j < 63 ? s_off[w][j + 1] : E;
            const int32_t x = s_lab[w][k];
            uint32_t c = 0;
            for (int32_t q = so; q < se; ++q) c += (s_lab[w][q] == x);
            atomicMax(&s_best[w][j], static_cast<unsigned long long>(pack(c, x)));
        }
        wave_lds_sync();
        if (i < count) out[v] = unpack_label(s_best[w][lane]);
        (void)off;
        wave_lds_sync();  // the next tile's gather overwrites s_lab and s_best
    }
}

// Counts the labels seg[0, d) (d <= 256) in the wave's 512-slot LDS table
// and returns the packed winner in every lane.
__device__ __forceinline__ uint64_t wave_hash_best(const int32_t* seg, int32_t d, int32_t* keys, uint32_t* cnts);

// 32 < d <= 256: one wave per vertex; its labels are loaded together into
// LDS and counted in the wave's hash table.
__global__ void __launch_bounds__(kBT) k_cdlp_wave(const int32_t* __restrict__ list, int32_t count,
                                                   const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                                   const int32_t* __restrict__ lab, int32_t* __restrict__ out) {
    __shared__ int32_t s_lab[kBT / 64][kWave];
    __shared__ int32_t s_keys[kBT / 64][2 * kWave];
    __shared__ uint32_t s_cnts[kBT / 64][2 * kWave];
    const int lane = lane_id(), w = threadIdx.x >> 6;
    for (int64_t i = (static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x) >> 6; i < count;
         i += (static_cast<int64_t>(gridDim.x) * kBT) >> 6) {
        const int32_t v = list[i];
        const int32_t b = rowptr[v], d = rowptr[v + 1] - b;
        int32_t cv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) cv[u] = col[b + min(u * 64 + lane, d - 1)];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int32_t x = lab[cv[u]];
            if (u * 64 + lane < d) s_lab[w][u * 64 + lane] = x;
        }
        wave_lds_sync();
        const uint64_t best = wave_hash_best(s_lab[w], d, s_keys[w], s_cnts[w]);
        if (lane == 0) out[v] = unpack_label(best);
        wave_lds_sync();
    }
}

// Inserts one label into an open-addressing table of `mask + 1` slots
// (keys kEmpty when free) and adds `c` to its count.  Returns false when the
// table has no slot left for it (only a capped merge table can fill up).
__device__ __forceinline__ bool table_add(int32_t* keys, uint32_t* cnts, uint32_t mask, int32_t x, uint32_t c) {
    uint32_t h = hash_slot(x, mask);
    for (uint32_t probe = 0; probe <= mask; ++probe) {
        const int32_t prev = atomicCAS(&keys[h], kEmpty, x);
        if (prev == kEmpty || prev == x) {
            atomicAdd(&cnts[h], c);
            return true;
        }
        h = (h + 1) & mask;
    }
    return false;
}

// table_add for a whole wave (every lane calls it, `valid` marks the lanes
// with a label): the two labels held by the first pending lanes are added
// once each with the number of lanes holding them, the rest lane by lane.
// Once labels converge most of a wave shares one or two of them, and their
// LDS atomics would otherwise serialise on one slot.  Returns false in the
// lanes whose add found the table full.
__device__ __forceinline__ bool wave_table_add(int32_t* keys, uint32_t* cnts, uint32_t mask, int32_t x, bool valid) {
    const int lane = lane_id();
    uint64_t pending = __ballot(valid);
    bool ok = true;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        if (!pending) break;
        const int leader = __ffsll(static_cast<unsigned long long>(pending)) - 1;
        const int32_t lx = __shfl(x, leader);
        const uint64_t same = __ballot(valid && x == lx) & pending;
        if (lane == leader) ok = table_add(keys, cnts, mask, lx, static_cast<uint32_t>(__popcll(same)));
        pending &= ~same;
    }
    if ((pending >> lane) & 1) ok = table_add(keys, cnts, mask, x, 1u) && ok;
    return ok;
}

__device__ __forceinline__ uint64_t wave_hash_best(const int32_t* seg, int32_t d, int32_t* keys, uint32_t* cnts) {
    const int lane = lane_id();
    uint32_t slots = 64;
    while (slots < 2u * static_cast<uint32_t>(d)) slots <<= 1;  // <= 512 for d <= 256
    for (uint32_t j = lane; j < slots; j += 64) {
        keys[j] = kEmpty;
        cnts[j] = 0;
    }
    wave_lds_sync();
    for (int32_t j0 = 0; j0 < d; j0 += 64) {
        const int32_t j = j0 + lane;
        wave_table_add(keys, cnts, slots - 1, j < d ? seg[j] : 0, j < d);
    }
    wave_lds_sync();
    uint64_t best = 0;
    for (uint32_t j = lane; j < slots; j += 64) {
        const int32_t k = keys[j];
        if (k != kEmpty) {
            const uint64_t key = pack(cnts[j], k);
            best = key > best ? key : best;
        }
    }
    best = wave_max(best);
    wave_lds_sync();  // the next call clears the table
    return best;
}

// Packed winner of the first `slots` table entries, reduced over the block
// (valid in every thread).  BT = threads per block.
template <int BT>
__device__ __forceinline__ uint64_t table_best(const int32_t* keys, const uint32_t* cnts, uint32_t slots,
                                               uint64_t* s_red) {
    uint64_t best = 0;
    for (uint32_t j = threadIdx.x; j < slots; j += BT) {
        const int32_t k = keys[j];
        if (k != kEmpty) {
            const uint64_t key = pack(cnts[j], k);
            best = key > best ? key : best;
        }
    }
    best = wave_max(best);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = best;
    __syncthreads();
    uint64_t t = 0;
    for (int w = 0; w < BT / 64; ++w) t = s_red[w] > t ? s_red[w] : t;
    __syncthreads();
    return t;
}

// Partition of `label` among `parts` (a multiply-shift hash; `mult` odd).
__device__ __forceinline__ uint32_t hash_part(int32_t label, uint32_t parts, uint32_t mult) {
    return static_cast<uint32_t>((static_cast<uint64_t>(static_cast<uint32_t>(label) * mult) * parts) >> 32);
}

__host__ __device__ __forceinline__ uint32_t pow2_at_least(uint32_t x) {
    uint32_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

// One block per vertex (grid-stride over the class list); the table uses the
// power of two >= 2d of its SLOTS slots, so only that much is cleared.  A
// vertex with more distinct labels than the table holds (only possible when
// 2d > SLOTS) is counted in 2, 4, ... sub-passes, each taking the labels of
// one sub-hash value and gathering its labels again.
template <int BT, int SLOTS>
__global__ void __launch_bounds__(BT) k_cdlp_lds(const int32_t* __restrict__ list, int32_t count,
                                                 const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                                 const int32_t* __restrict__ lab, int32_t* __restrict__ out) {
    __shared__ int32_t keys[SLOTS];
    __shared__ uint32_t cnts[SLOTS];
    __shared__ uint64_t s_red[BT / 64];
    __shared__ int s_full;
    for (int64_t i = blockIdx.x; i < count; i += gridDim.x) {
        const int32_t v = list[i];
        const int32_t b = rowptr[v], d = rowptr[v + 1] - b;
        const uint32_t want = pow2_at_least(2u * static_cast<uint32_t>(d));
        const uint32_t slots = want < static_cast<uint32_t>(SLOTS) ? want : static_cast<uint32_t>(SLOTS);
        uint64_t best = 0;
        for (uint32_t nsub = 1;; nsub <<= 1) {
            bool full = false;
            uint64_t cand = 0;
            for (uint32_t sub = 0; sub < nsub && !full; ++sub) {
                for (uint32_t j = threadIdx.x; j < slots; j += BT) {
                    keys[j] = kEmpty;
                    cnts[j] = 0;
                }
                if (threadIdx.x == 0) s_full = 0;
                __syncthreads();
                // whole waves (wave_table_add); each lane's kLdsLoads edges
                // load together (d >= 1: lanes past the end read the last edge)
                for (int j0 = threadIdx.x & ~63; j0 < d; j0 += BT * kLdsLoads) {
                    int32_t u[kLdsLoads], x[kLdsLoads];
#pragma unroll
                    for (int k = 0; k < kLdsLoads; ++k) {
                        const int j = j0 + k * BT + lane_id();
                        u[k] = col[b + (j < d ? j : d - 1)];
                    }
#pragma unroll
                    for (int k = 0; k < kLdsLoads; ++k) x[k] = lab[u[k]];
#pragma unroll
                    for (int k = 0; k < kLdsLoads; ++k) {
                        const int j = j0 + k * BT + lane_id();
                        const bool mine = j < d && (nsub == 1 || hash_part(x[k], nsub, 0x27D4EB2Fu) == sub);
                        if (!wave_table_add(keys, cnts, slots - 1, x[k], mine)) s_full = 1;
                    }
                }
                __syncthreads();
                full = s_full != 0;
                if (!full) {
                    const uint64_t t = table_best<BT>(keys, cnts, slots, s_red);
                    cand = t > cand ? t : cand;
                } else {
                    __syncthreads();  // every thread has read s_full before it is cleared
                }
            }
            if (!full) {
                best = cand;
                break;
            }
        }
        if (threadIdx.x == 0) out[v] = unpack_label(best);
    }
}

// Hub chunk c: count the labels of edges [beg, beg + kChunk) in an LDS table
// and write the distinct (label, count) pairs to pairs[c * kChunk ...]; the
// pair count goes to chunk_pairs[c] and is added to the hub's total.
__global__ void __launch_bounds__(kBigBT) k_cdlp_hub_count(
    const int32_t* __restrict__ hubs, const int32_t* __restrict__ chunk_hub, const int32_t* __restrict__ chunk_beg,
    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col, const int32_t* __restrict__ lab,
    uint2* __restrict__ pairs, int32_t* __restrict__ chunk_pairs, int32_t* __restrict__ hub_pairs) {
    __shared__ int32_t keys[kBigSlots];
    __shared__ uint32_t cnts[kBigSlots];
    __shared__ int32_t s_n;
    const int32_t c = blockIdx.x, h = chunk_hub[c], v = hubs[h];
    const int32_t beg = chunk_beg[c], end = min(beg + kChunk, rowptr[v + 1]);
    const uint32_t slots = pow2_at_least(2u * static_cast<uint32_t>(end - beg));
    for (uint32_t j = threadIdx.x; j < slots; j += kBigBT) {
        keys[j] = kEmpty;
        cnts[j] = 0;
    }
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    // whole waves (wave_table_add); the chunk's kHubLoads edges per lane load
    // together (lanes past the end read the chunk's last edge)
    {
        int32_t u[kHubLoads], x[kHubLoads];
        const int32_t j0 = beg + (threadIdx.x & ~63);
#pragma unroll
        for (int k = 0; k < kHubLoads; ++k) {
            const int32_t j = j0 + k * kBigBT + lane_id();
            u[k] = col[j < end ? j : end - 1];
        }
#pragma unroll
        for (int k = 0; k < kHubLoads; ++k) x[k] = lab[u[k]];
#pragma unroll
        for (int k = 0; k < kHubLoads; ++k) {
            const int32_t j = j0 + k * kBigBT + lane_id();
            wave_table_add(keys, cnts, slots - 1, x[k], j < end);
        }
    }
    __syncthreads();
    uint2* out = pairs + static_cast<int64_t>(c) * kChunk;
    for (uint32_t j = threadIdx.x; j < slots; j += kBigBT) {
        const int32_t k = keys[j];
        const int slot = wave_append(&s_n, k != kEmpty);
        if (k != kEmpty) out[slot] = make_uint2(static_cast<uint32_t>(k), cnts[j]);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        chunk_pairs[c] = s_n;
        atomicAdd(&hub_pairs[h], s_n);
    }
}

// Merge items per hub: ceil(pairs / kMergeTarget) (at least one), as an
// exclusive scan into item_off[0..hubs]; clears the pair totals.  One block.
__global__ void __launch_bounds__(1024) k_cdlp_hub_plan(int32_t* __restrict__ hub_pairs, int32_t hubs,
                                                        int32_t* __restrict__ item_off) {
    __shared__ int32_t s_wt[16];
    __shared__ int32_t s_base;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) s_base = 0;
    __syncthreads();
    for (int32_t h0 = 0; h0 < hubs; h0 += 1024) {
        const int32_t h = h0 + tid;
        int32_t p = 0;
        if (h < hubs) {
            p = max(1, (hub_pairs[h] + kMergeTarget - 1) / kMergeTarget);
            hub_pairs[h] = 0;
        }
        int32_t incl = p;
        for (int off = 1; off < 64; off <<= 1) {
            const int32_t y = __shfl_up(incl, off);
            if (lane >= off) incl += y;
        }
        if (lane == 63) s_wt[wave] = incl;
        __syncthreads();
        int32_t wb = s_base;
        for (int w = 0; w < wave; ++w) wb += s_wt[w];
        if (h < hubs) item_off[h] = wb + incl - p;
        __syncthreads();
        if (tid == 1023) s_base = wb + incl;
        __syncthreads();
    }
    if (tid == 0) item_off[hubs] = s_base;
}

// Merge item i = (hub h, partition p of P): the pairs of h's chunks whose
// label hashes to p are added up in an LDS table and the best goes to
// hub_best[h] by atomicMax (max is order-free, so the result does not depend
// on which item lands first).  If the table fills up, the partition is redone
// in 2, 4, ... sub-passes (each taking the labels of one sub-hash value)
// until every sub-pass fits.  Persistent blocks walk the items.  cap (a power
// of two) bounds the table; a cap of 64 is a test setting that forces the
// sub-passes.

__global__ void __launch_bounds__(kBigBT) k_cdlp_hub_merge(
    const int32_t* __restrict__ item_off, int32_t hubs, const int32_t* __restrict__ hub_chunk0,
    const uint2* __restrict__ pairs, const int32_t* __restrict__ chunk_pairs, uint32_t cap,
    unsigned long long* __restrict__ hub_best) {
    __shared__ int32_t keys[kBigSlots];
    __shared__ uint32_t cnts[kBigSlots];
    __shared__ int32_t s_pre[kMaxHubChunks + 1];  // pair prefix over the hub's chunks
    __shared__ uint64_t s_red[kBigBT / 64];
    __shared__ int s_full;
    const int tid = threadIdx.x;
    const int32_t items = item_off[hubs];
    for (int32_t i = blockIdx.x; i < items; i += gridDim.x) {
        int32_t lo = 0, hi = hubs;  // last h with item_off[h] <= i (item_off[hubs] = items > i)
        while (hi - lo > 1) {
            const int32_t mid = (lo + hi) >> 1;
            if (item_off[mid] <= i) lo = mid;
            else hi = mid;
        }
        const int32_t h = lo, P = item_off[h + 1] - item_off[h], p = i - item_off[h];
        const int32_t c0 = hub_chunk0[h], nc = hub_chunk0[h + 1] - c0;
        // prefix of the chunks' pair counts (nc <= kMaxHubChunks, checked by
        // the layout): loaded in parallel, summed by one thread in LDS
        for (int32_t k = tid; k < nc; k += kBigBT) s_pre[k + 1] = chunk_pairs[c0 + k];
        __syncthreads();
        if (tid == 0) {
            s_pre[0] = 0;
            for (int32_t k = 1; k <= nc; ++k) s_pre[k] += s_pre[k - 1];
        }
        __syncthreads();
        const int32_t total = s_pre[nc];
        // a table of four slots per pair this partition should get (at
        // least 64, at most cap); a partition that gets more distinct
        // labels than that takes the sub-pass path below
        const uint32_t want = pow2_at_least(4u * static_cast<uint32_t>((total + P - 1) / P));
        const uint32_t slots = want < 64u ? 64u : (want > cap ? cap : want);
        uint64_t best = 0;
        for (uint32_t nsub = 1;; nsub <<= 1) {
            bool full = false;
            uint64_t cand = 0;
            for (uint32_t sub = 0; sub < nsub && !full; ++sub) {
                for (uint32_t j = tid; j < slots; j += kBigBT) {
                    keys[j] = kEmpty;
                    cnts[j] = 0;
                }
                if (tid == 0) s_full = 0;
                __syncthreads();
                for (int32_t k = tid; k < total; k += kBigBT) {
                    int32_t a = 0, b = nc;  // chunk of pair k: last a with s_pre[a] <= k
                    while (b - a > 1) {
                        const int32_t mid = (a + b) >> 1;
                        if (s_pre[mid] <= k) a = mid;
                        else b = mid;
                    }
                    const uint2 pr = pairs[static_cast<int64_t>(c0 + a) * kChunk + (k - s_pre[a])];
                    const int32_t x = static_cast<int32_t>(pr.x);
                    if (hash_part(x, P, 0xC2B2AE35u) == static_cast<uint32_t>(p) &&
                        hash_part(x, nsub, 0x27D4EB2Fu) == sub && !table_add(keys, cnts, slots - 1, x, pr.y))
                        s_full = 1;
                }
                __syncthreads();
                full = s_full != 0;
                if (!full) {
                    const uint64_t t = table_best<kBigBT>(keys, cnts, slots, s_red);
                    cand = t > cand ? t : cand;
                } else {
                    __syncthreads();
                }
            }
            if (!full) {
                best = cand;
                break;
            }
        }
        if (tid == 0 && best) atomicMax(&hub_best[h], static_cast<unsigned long long>(best));
        __syncthreads();
    }
}

// out[hub] = the merged winner; the slot is cleared for the next iteration.
__global__ void k_cdlp_hub_final(const int32_t* __restrict__ hubs, int32_t count,
                                 unsigned long long* __restrict__ hub_best, int32_t* __restrict__ out) {
    for (int64_t h = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; h < count;
         h += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        out[hubs[h]] = unpack_label(hub_best[h]);
        hub_best[h] = 0ull;
    }
}

// First iteration.  Labels start as vertex ids, and a vertex's all-list is
// its in-list then its out-list, each sorted and duplicate-free, so a label
// is counted twice exactly when that neighbour is both an in- and an
// out-neighbour and once otherwise.  The mode is therefore the smallest
// neighbour present in both lists if there is one (count 2), else the
// smallest neighbour (all counts 1); a vertex without neighbours keeps its
// id.  Undirected graphs (all-list = out-list, in_rowptr == nullptr) take
// the smallest neighbour.  Same result as the counting kernels, no hashing.
__device__ __forceinline__ bool contains_sorted(const int32_t* __restrict__ a, int32_t len, int32_t x) {
    int32_t lo = 0, hi = len;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        const int32_t y = a[mid];
        if (y == x) return true;
        if (y < x) lo = mid + 1;
        else hi = mid;
    }
    return false;
}

__device__ __forceinline__ int32_t first_fallback(const int32_t* __restrict__ col, int32_t b, int32_t e, int32_t di,
                                                  int32_t v) {
    if (e == b) return v;
    if (di == 0 || di == e - b) return col[b];  // one list: its first element
    return min(col[b], col[b + di]);
}

__global__ void __launch_bounds__(kBT) k_cdlp_first_small(const int32_t* __restrict__ list, int32_t count,
                                                          const int32_t* __restrict__ rowptr,
                                                          const int32_t* __restrict__ col,
                                                          const int32_t* __restrict__ in_rowptr,
                                                          int32_t* __restrict__ out) {
    for (int64_t k = static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x; k < count;
         k += static_cast<int64_t>(gridDim.x) * kBT) {
        const int32_t v = list[k];
        const int32_t b = rowptr[v], e = rowptr[v + 1];
        const int32_t di = in_rowptr ? in_rowptr[v + 1] - in_rowptr[v] : 0;
        int32_t best = -1;
        for (int32_t i = b, j = b + di; i < b + di && j < e;) {  // merge: first common = smallest
            const int32_t x = col[i], y = col[j];
            if (x == y) {
                best = x;
                break;
            }
            if (x < y) ++i;
            else ++j;
        }
        out[v] = best >= 0 ? best : first_fallback(col, b, e, di, v);
    }
}

// One wave per vertex; in-list scanned in 64-element steps, each lane
// binary-searching its element in the out-list; the first hit is the smallest.
__global__ void __launch_bounds__(kBT) k_cdlp_first_wave(const int32_t* __restrict__ list, int32_t count,
                                                         const int32_t* __restrict__ rowptr,
                                                         const int32_t* __restrict__ col,
                                                         const int32_t* __restrict__ in_rowptr,
                                                         int32_t* __restrict__ out) {
    const int lane = lane_id();
    const int64_t wave = (static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x) >> 6;
    const int64_t waves = (static_cast<int64_t>(gridDim.x) * kBT) >> 6;
    for (int64_t k = wave; k < count; k += waves) {
        const int32_t v = list[k];
        const int32_t b = rowptr[v], e = rowptr[v + 1];
        const int32_t di = in_rowptr ? in_rowptr[v + 1] - in_rowptr[v] : 0;
        int32_t best = -1;
        for (int32_t c0 = 0; c0 < di && di < e - b; c0 += 64) {
            const int32_t i = c0 + lane;
            const bool hit = i < di && contains_sorted(col + b + di, e - b - di, col[b + i]);
            const uint64_t m = __ballot(hit);
            if (m) {
                best = col[b + c0 + __ffsll(static_cast<unsigned long long>(m)) - 1];
                break;
            }
        }
        if (lane == 0) out[v] = best >= 0 ? best : first_fallback(col, b, e, di, v);
    }
}

// One workgroup per vertex of more than 2048 neighbours; the in-list in
// 256-element steps.
__global__ void __launch_bounds__(kBT) k_cdlp_first_block(const int32_t* __restrict__ list, int32_t count,
                                                          const int32_t* __restrict__ rowptr,
                                                          const int32_t* __restrict__ col,
                                                          const int32_t* __restrict__ in_rowptr,
                                                          int32_t* __restrict__ out) {
    __shared__ int32_t s_best;
    for (int64_t k = blockIdx.x; k < count; k += gridDim.x) {
        const int32_t v = list[k];
        const int32_t b = rowptr[v], e = rowptr[v + 1];
        const int32_t di = in_rowptr ? in_rowptr[v + 1] - in_rowptr[v] : 0;
        if (threadIdx.x == 0) s_best = INT32_MAX;
        __syncthreads();
        for (int32_t c0 = 0; c0 < di && di < e - b; c0 += kBT) {
            const int32_t i = c0 + static_cast<int32_t>(threadIdx.x);
            if (i < di) {
                const int32_t x = col[b + i];
                if (contains_sorted(col + b + di, e - b - di, x)) atomicMin(&s_best, x);
            }
            __syncthreads();
            const bool done = s_best != INT32_MAX;
            __syncthreads();
            if (done) break;
        }
        if (threadIdx.x == 0) out[v] = s_best != INT32_MAX ? s_best : first_fallback(col, b, e, di, v);
        __syncthreads();
    }
}

unsigned grid_cap(int64_t blocks) {
    return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(blocks, 65535)));
}

// Rank order: all-degree descending, ties by vertex id (a stable sort of
// these keys).
__global__ void k_cdlp_rank_keys(const int32_t* __restrict__ rp, int32_t n, uint32_t* __restrict__ key,
                                 int32_t* __restrict__ val) {
    for (int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < n;
         v += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        key[v] = 0xFFFFFFFFu - static_cast<uint32_t>(rp[v + 1] - rp[v]);
        val[v] = static_cast<int32_t>(v);
    }
}

// Position of each class in rank order: hubs first, then classes 3, 2, 1, 0
// and the vertices without neighbours (degrees descend along the ranks).
__host__ __device__ constexpr int class_order(int c) { return c == 5 ? 5 : 4 - c; }

__device__ __forceinline__ int degree_class(int32_t d) {
    return d == 0 ? 5 : d <= kSmall ? 0 : d <= kWave ? 1 : d <= kLds ? 2 : d <= kBig ? 3 : 4;
}

// Keys of a stable sort of the vertex ids by class position.
__global__ void k_cdlp_first_keys(const int32_t* __restrict__ rp, int32_t n, uint32_t* __restrict__ key,
                                  int32_t* __restrict__ val) {
    for (int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < n;
         v += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        key[v] = static_cast<uint32_t>(class_order(degree_class(rp[v + 1] - rp[v])));
        val[v] = static_cast<int32_t>(v);
    }
}

// rank[perm[r]] = r; deg_r[r] = all-degree of rank r; okey[r] = the class
// position of rank r (non-decreasing along the ranks); iota[r] = r.
__global__ void k_cdlp_rank_rows(const int32_t* __restrict__ perm, const int32_t* __restrict__ rp, int32_t n,
                                 int32_t* __restrict__ rank, int32_t* __restrict__ deg_r, uint32_t* __restrict__ okey,
                                 int32_t* __restrict__ iota) {
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; r < n;
         r += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int32_t v = perm[r];
        const int32_t d = rp[v + 1] - rp[v];
        rank[v] = static_cast<int32_t>(r);
        deg_r[r] = d;
        okey[r] = static_cast<uint32_t>(class_order(degree_class(d)));
        iota[r] = static_cast<int32_t>(r);
    }
}

// off[k] = first rank of class position k in the non-decreasing keys; off[6] = n.
__global__ void k_cdlp_class_offsets(const uint32_t* __restrict__ key, int32_t n, int32_t* __restrict__ off) {
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e <= n;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t prev = e > 0 ? static_cast<int64_t>(key[e - 1]) : -1;
        const int64_t cur = e < n ? static_cast<int64_t>(key[e]) : 6;
        for (int64_t k = prev + 1; k <= cur; ++k) off[k] = static_cast<int32_t>(e);
    }
}

// The all-CSR in rank space, edge-balanced: a block takes kColTile
// consecutive positions of col_r, finds their rows (ranks) by binary search
// in an LDS copy of the tile's row starts, and writes rank[col[...]] of the
// same edge of the vertex's own row.  Rows keep their edge order here; a
// segmented sort orders each afterwards.  Rows without edges come last in
// rank order, so every row a position falls in has edges.
constexpr int kColBT = 256, kColItems = 8, kColTile = kColBT * kColItems;

__global__ void __launch_bounds__(kColBT) k_cdlp_rank_cols(const int32_t* __restrict__ perm,
                                                           const int32_t* __restrict__ rp,
                                                           const int32_t* __restrict__ col,
                                                           const int32_t* __restrict__ rank,
                                                           const int32_t* __restrict__ rowptr_r, int32_t n, int64_t m,
                                                           int32_t* __restrict__ col_r) {
    __shared__ int32_t s_start[kColTile], s_src[kColTile];
    __shared__ int32_t s_r0, s_span;
    for (int64_t tile = blockIdx.x; tile * kColTile < m; tile += gridDim.x) {
        const int64_t t0 = tile * kColTile;
        const int64_t t1 = std::min<int64_t>(t0 + kColTile, m);
        if (threadIdx.x == 0) {
            auto row_of = [&](int64_t pos) {  // last r with rowptr_r[r] <= pos
                int32_t lo = 0, hi = n;
                while (lo < hi) {
                    const int32_t mid = (lo + hi + 1) >> 1;
                    if (rowptr_r[mid] <= pos) lo = mid;
                    else hi = mid - 1;
                }
                return lo;
            };
            const int32_t r0 = row_of(t0), r1 = row_of(t1 - 1);
            s_r0 = r0;
            s_span = r1 - r0 + 1;  // <= kColTile: each row in it has an edge in the tile
        }
        __syncthreads();
        const int32_t r0 = s_r0, span = s_span;
        for (int j = threadIdx.x; j < span; j += kColBT) {
            s_start[j] = rowptr_r[r0 + j];
            s_src[j] = rp[perm[r0 + j]];
        }
        __syncthreads();
        int32_t e[kColItems];
#pragma unroll
        for (int k = 0; k < kColItems; ++k) {
            const int64_t pos = t0 + k * kColBT + threadIdx.x;
            e[k] = 0;
            if (pos < t1) {
                int32_t lo = 0, hi = span - 1;  // last j with s_start[j] <= pos
                while (lo < hi) {
                    const int32_t mid = (lo + hi + 1) >> 1;
                    if (s_start[mid] <= pos) lo = mid;
                    else hi = mid - 1;
                }
                e[k] = s_src[lo] + static_cast<int32_t>(pos - s_start[lo]);
            }
        }
        int32_t u[kColItems];
#pragma unroll
        for (int k = 0; k < kColItems; ++k) u[k] = col[e[k]];  // m > 0: e = 0 is an edge
#pragma unroll
        for (int k = 0; k < kColItems; ++k) {
            const int64_t pos = t0 + k * kColBT + threadIdx.x;
            if (pos < t1) col_r[pos] = rank[u[k]];
        }
        __syncthreads();  // the next tile rewrites the LDS copies
    }
}

// Labels by vertex id -> rank space (a and b: the pair of label buffers).
__global__ void k_cdlp_to_rank(const int32_t* __restrict__ perm, const int32_t* __restrict__ lo, int32_t n,
                               int32_t* __restrict__ a, int32_t* __restrict__ b) {
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; r < n;
         r += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int32_t x = lo[perm[r]];
        a[r] = x;
        b[r] = x;
    }
}

// Rank space -> labels by vertex id.
__global__ void k_cdlp_from_rank(const int32_t* __restrict__ perm, const int32_t* __restrict__ a, int32_t n,
                                 int32_t* __restrict__ lo) {
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; r < n;
         r += static_cast<int64_t>(gridDim.x) * blockDim.x)
        lo[perm[r]] = a[r];
}

// Chunks per hub.
__global__ void k_cdlp_hub_nch(const int32_t* __restrict__ rp, const int32_t* __restrict__ hubs, int32_t h,
                               int32_t* __restrict__ nch) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < h;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int32_t v = hubs[i];
        nch[i] = (rp[v + 1] - rp[v] + kChunk - 1) / kChunk;
    }
}

__global__ void k_cdlp_hub_chunks(const int32_t* __restrict__ rp, const int32_t* __restrict__ hubs, int32_t h,
                                  const int32_t* __restrict__ choff, int32_t* __restrict__ cv,
                                  int32_t* __restrict__ cb) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < h;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int32_t v = hubs[i];
        int32_t o = choff[i];
        for (int32_t e = rp[v]; e < rp[v + 1]; e += kChunk, ++o) {
            cv[o] = static_cast<int32_t>(i);
            cb[o] = e;
        }
    }
}

// The rank order from a stable radix sort of the degree keys, the all-CSR in
// rank space (edge-balanced gather, then a segmented radix sort of each
// row), the class ranges, and the hubs' chunks from a scan; all on the GPU.
int build_cdlp_layout(ghet_graph* g, hipStream_t s) {
    CDLPLayout& L = g->cdlp;
    if (L.built) return GHET_OK;
    const int32_t n = g->n;
    const int32_t* rp = g->all.rowptr;
    const int64_t m = g->all.m;
    if (m >= (int64_t(1) << 31)) return set_error(GHET_ERR_UNSUPPORTED, "CDLP: %lld all-edges", static_cast<long long>(m));
    DevTemp tmp;
    uint32_t* ka = tmp.get<uint32_t>(size_t(n));
    uint32_t* kb = tmp.get<uint32_t>(size_t(n));
    int32_t* va = tmp.get<int32_t>(size_t(n));
    int32_t* vb = tmp.get<int32_t>(size_t(n));
    int32_t* rank = tmp.get<int32_t>(size_t(n));
    int32_t* deg_r = tmp.get<int32_t>(size_t(n));
    int32_t* off = tmp.get<int32_t>(8);
    int32_t* aux = tmp.get<int32_t>(size_t(scan_aux_ints(n)));
    int32_t* col_tmp = tmp.get<int32_t>(size_t(std::max<int64_t>(m, 1)));
    hipError_t e = tmp.err;
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    // the rank-space arrays live as long as the graph
    const size_t rbytes = 3 * al(4 * size_t(n)) + al(4 * (size_t(n) + 1)) + al(4 * size_t(std::max<int64_t>(m, 1)));
    void* rmem = nullptr;
    if (e == hipSuccess) e = hipMalloc(&rmem, rbytes);
    int32_t *perm = nullptr, *iota = nullptr, *rowptr_r = nullptr, *col_r = nullptr, *first_ids = nullptr;
    if (e == hipSuccess) {
        char* q = static_cast<char*>(rmem);
        perm = reinterpret_cast<int32_t*>(q); q += al(4 * size_t(n));
        first_ids = reinterpret_cast<int32_t*>(q); q += al(4 * size_t(n));
        iota = reinterpret_cast<int32_t*>(q); q += al(4 * size_t(n));
        rowptr_r = reinterpret_cast<int32_t*>(q); q += al(4 * (size_t(n) + 1));
        col_r = reinterpret_cast<int32_t*>(q);
    }
    const unsigned gn = grid_cap(ceil_div(n, 256));
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_cdlp_rank_keys, dim3(gn), dim3(256), 0, s, rp, n, ka, va);
        e = hipGetLastError();
    }
    rocprim::double_buffer<uint32_t> kd(ka, kb);
    rocprim::double_buffer<int32_t> vd(va, vb);
    size_t sb = 0;
    void* sort_tmp = nullptr;
    if (e == hipSuccess) e = rocprim::radix_sort_pairs(nullptr, sb, kd, vd, static_cast<unsigned>(n), 0, 32, s);
    if (e == hipSuccess) sort_tmp = tmp.get<char>(sb), e = tmp.err;
    if (e == hipSuccess) e = rocprim::radix_sort_pairs(sort_tmp, sb, kd, vd, static_cast<unsigned>(n), 0, 32, s);
    if (e == hipSuccess) e = hipMemcpyAsync(perm, vd.current(), 4 * size_t(n), hipMemcpyDeviceToDevice, s);
    // kb's keys are no longer needed: it takes the class positions
    uint32_t* okey = kd.current() == ka ? kb : ka;
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_cdlp_rank_rows, dim3(gn), dim3(256), 0, s, perm, rp, n, rank, deg_r, okey, iota);
        e = hipGetLastError();
    }
    if (e == hipSuccess && exclusive_scan_i32(deg_r, rowptr_r, n, aux, s) != GHET_OK) e = hipErrorUnknown;
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_cdlp_class_offsets, dim3(grid_cap(ceil_div(int64_t(n) + 1, 256))), dim3(256), 0, s, okey,
                           n, off);
        e = hipGetLastError();
    }
    if (e == hipSuccess && m > 0) {
        hipLaunchKernelGGL(k_cdlp_rank_cols, dim3(grid_cap(ceil_div(m, kColTile))), dim3(kColBT), 0, s, perm, rp,
                           g->all.col, rank, rowptr_r, n, m, col_tmp);
        e = hipGetLastError();
    }
    // each row's ranks ascending: consecutive lanes of a counting kernel
    // then read neighbouring labels, the hubs' first
    int end_bit = 1;
    while (end_bit < 31 && (int64_t(1) << end_bit) < n) ++end_bit;
    // the closed-form first iteration reads the all-CSR by vertex id: each
    // class's vertices in id order (a stable sort of the ids by class
    // position), so consecutive waves read neighbouring rows
    if (e == hipSuccess) {
        uint32_t* fk0 = tmp.get<uint32_t>(size_t(n));
        uint32_t* fk1 = tmp.get<uint32_t>(size_t(n));
        int32_t* fv0 = tmp.get<int32_t>(size_t(n));
        e = tmp.err;
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_cdlp_first_keys, dim3(gn), dim3(256), 0, s, rp, n, fk0, fv0);
            e = hipGetLastError();
        }
        rocprim::double_buffer<uint32_t> fkd(fk0, fk1);
        rocprim::double_buffer<int32_t> fvd(fv0, first_ids);
        size_t fb = 0;
        if (e == hipSuccess) e = rocprim::radix_sort_pairs(nullptr, fb, fkd, fvd, static_cast<unsigned>(n), 0, 3, s);
        void* ftmp = nullptr;
        if (e == hipSuccess) ftmp = tmp.get<char>(fb), e = tmp.err;
        if (e == hipSuccess) e = rocprim::radix_sort_pairs(ftmp, fb, fkd, fvd, static_cast<unsigned>(n), 0, 3, s);
        if (e == hipSuccess && fvd.current() != first_ids)
            e = hipMemcpyAsync(first_ids, fvd.current(), 4 * size_t(n), hipMemcpyDeviceToDevice, s);
    }
    if (e == hipSuccess && m > 0) {
        size_t seg_b = 0;
        e = rocprim::segmented_radix_sort_keys(nullptr, seg_b, col_tmp, col_r, static_cast<unsigned>(m),
                                               static_cast<unsigned>(n), rowptr_r, rowptr_r + 1, 0, end_bit, s);
        void* seg_tmp = nullptr;
        if (e == hipSuccess) seg_tmp = tmp.get<char>(seg_b), e = tmp.err;
        if (e == hipSuccess)
            e = rocprim::segmented_radix_sort_keys(seg_tmp, seg_b, col_tmp, col_r, static_cast<unsigned>(m),
                                                   static_cast<unsigned>(n), rowptr_r, rowptr_r + 1, 0, end_bit, s);
    }
    int32_t h_off[7] = {};
    if (e == hipSuccess) e = hipMemcpyAsync(h_off, off, sizeof(h_off), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    int32_t cbeg[6] = {}, ccnt[6] = {};
    for (int c = 0; c < 6; ++c) {
        const int k = class_order(c);
        cbeg[c] = h_off[k];
        ccnt[c] = (k + 1 <= 5 ? h_off[k + 1] : n) - h_off[k];
    }
    const int32_t nh = ccnt[4];
    const int32_t* hubs = iota + cbeg[4];
    int32_t* nch = tmp.get<int32_t>(size_t(nh));
    int32_t* choff = tmp.get<int32_t>(size_t(nh) + 1);
    int32_t* aux32 = tmp.get<int32_t>(size_t(scan_aux_ints(nh)));
    if (e == hipSuccess) e = tmp.err;
    if (e == hipSuccess && nh > 0) {
        hipLaunchKernelGGL(k_cdlp_hub_nch, dim3(grid_cap(ceil_div(nh, 256))), dim3(256), 0, s, rowptr_r, hubs, nh, nch);
        e = hipGetLastError();
    }
    if (e == hipSuccess && exclusive_scan_i32(nch, choff, nh, aux32, s) != GHET_OK) e = hipErrorUnknown;
    std::vector<int32_t> h_nch(static_cast<size_t>(nh));
    int32_t chunks = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&chunks, choff + nh, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && nh > 0) e = hipMemcpyAsync(h_nch.data(), nch, 4 * size_t(nh), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    const int32_t max_nch = nh > 0 ? *std::max_element(h_nch.begin(), h_nch.end()) : 0;
    if (e == hipSuccess && max_nch > kMaxHubChunks) {
        hipFree(rmem);
        return set_error(GHET_ERR_UNSUPPORTED, "CDLP: a vertex of more than %d neighbours", kMaxHubChunks * kChunk);
    }
    const size_t bytes = 2 * al(4 * size_t(chunks)) + al(4 * (size_t(nh) + 1)) + al(8 * size_t(chunks) * kChunk) +
                         al(4 * size_t(chunks)) + al(4 * size_t(nh)) + al(4 * (size_t(nh) + 1)) + al(8 * size_t(nh));
    if (e == hipSuccess) e = hipMalloc(&L.mem, std::max<size_t>(bytes, 256));
    if (e == hipSuccess) {
        char* q = static_cast<char*>(L.mem);
        auto take = [&](size_t b) { char* r = q; q += al(b); return r; };
        for (int c = 0; c < 6; ++c) {
            L.count[c] = ccnt[c];
            L.list[c] = iota + cbeg[c];
            L.first[c] = first_ids + cbeg[c];
        }
        L.perm = perm;
        L.iota = iota;
        L.rowptr_r = rowptr_r;
        L.col_r = col_r;
        L.rmem = rmem;
        L.hub_chunks = chunks;
        L.max_hub_chunks = max_nch;
        L.chunk_hub = reinterpret_cast<int32_t*>(take(4 * size_t(chunks)));
        L.chunk_beg = reinterpret_cast<int32_t*>(take(4 * size_t(chunks)));
        L.hub_chunk0 = reinterpret_cast<int32_t*>(take(4 * (size_t(nh) + 1)));
        L.pairs = reinterpret_cast<uint2*>(take(8 * size_t(chunks) * kChunk));
        L.chunk_pairs = reinterpret_cast<int32_t*>(take(4 * size_t(chunks)));
        L.hub_pairs = reinterpret_cast<int32_t*>(take(4 * size_t(nh)));
        L.item_off = reinterpret_cast<int32_t*>(take(4 * (size_t(nh) + 1)));
        L.hub_best = reinterpret_cast<unsigned long long*>(take(8 * size_t(nh)));
        e = hipMemcpyAsync(L.hub_chunk0, choff, 4 * (size_t(nh) + 1), hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess && nh > 0) {
            hipLaunchKernelGGL(k_cdlp_hub_chunks, dim3(grid_cap(ceil_div(nh, 256))), dim3(256), 0, s, rowptr_r, hubs, nh,
                               choff, L.chunk_hub, L.chunk_beg);
            e = hipGetLastError();
        }
        // the plan clears hub_pairs after reading them and the final clears
        // hub_best, so they start at zero once here (fill kernels, no memsets)
        if (e == hipSuccess && nh > 0 && fill<int32_t>(L.hub_pairs, 0, nh, s) != GHET_OK) e = hipErrorUnknown;
        if (e == hipSuccess && nh > 0 && fill<uint64_t>(reinterpret_cast<uint64_t*>(L.hub_best), 0, nh, s) != GHET_OK)
            e = hipErrorUnknown;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        if (L.mem) hipFree(L.mem);
        hipFree(rmem);
        L = CDLPLayout{};
        return set_error(GHET_ERR_HIP, "CDLP layout: %s", hipGetErrorString(e));
    }
    L.built = true;
    return GHET_OK;
}

}  // namespace

int run_cdlp(ghet_graph* g, int32_t iters, int32_t* label_out, const ghet_opts* o, ghet_run_stats* st) {
    const int32_t n = g->n;
    if (iters < 0) return set_error(GHET_ERR_INVALID, "iterations %d < 0", iters);
    if (!g->all.present() || g->h_all_rowptr.empty())
        return set_error(GHET_ERR_INVALID, "CDLP needs the all-CSR (in-list ++ out-list)");
    if (n == 0) return GHET_OK;
    hipStream_t s = pick_stream(g, o);
    GHET_CHECK(build_cdlp_layout(g, s));
    CDLPLayout& L = g->cdlp;
    void* p;
    int32_t *la, *lb, *lo;  // la, lb: labels in rank space; lo: by vertex id
    GHET_CHECK(g->ws.get(WS_LABEL_A, 4 * size_t(n), &p)); la = static_cast<int32_t*>(p);
    GHET_CHECK(g->ws.get(WS_LABEL_B, 4 * size_t(n), &p)); lb = static_cast<int32_t*>(p);
    GHET_CHECK(g->ws.get(WS_HASH, 4 * size_t(n), &p)); lo = static_cast<int32_t*>(p);
    const bool timing = st != nullptr;
    hipEvent_t* kev = nullptr;
    if (timing) GHET_CHECK(kernel_events(g, 2 * std::max(iters, 1), &kev));
    // hub merge: persistent blocks (one per CU: the table is 128 KB of LDS);
    // opts reserved[0] bit 2 caps its table at 64 slots, a test setting that
    // makes the merge take its sub-pass path
    int cus = 0;
    GHET_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device));
    const int32_t merge_blocks = std::max(1, cus);
    const bool small_cap = o && o->size >= sizeof(ghet_opts) && (o->reserved[0] & 4);
    const uint32_t merge_cap = small_cap ? 64u : static_cast<uint32_t>(kBigSlots);
    const unsigned gn = grid_cap(ceil_div(n, 256));

    GHET_HIP(hipEventRecord(g->ev_begin, s));
    hipLaunchKernelGGL(k_label_init, dim3(gn), dim3(256), 0, s, lo, lo, n);
    // the identity-label first iteration has a closed form (see k_cdlp_first_*);
    // it needs the in-degrees to split each all-list, or an undirected graph
    const bool first_closed = !g->directed || (g->in.present() && g->in.rowptr != g->all.rowptr);
    const int32_t* in_rp = g->directed ? g->in.rowptr : nullptr;
    bool ranked = false;  // the current labels are in la (rank space), else in lo
    for (int32_t it = 0; it < iters; ++it) {
        if (timing) GHET_HIP(hipEventRecord(kev[2 * it], s));
        if (it == 0 && first_closed) {
            // on the all-CSR itself, by vertex id; vertices with no neighbours
            // are in no list and keep their id (lo holds ids)
            const int32_t* A = g->all.rowptr;
            const int32_t* C = g->all.col;
            if (L.count[0])
                hipLaunchKernelGGL(k_cdlp_first_small, dim3(grid_cap(ceil_div(L.count[0], kBT))), dim3(kBT), 0, s,
                                   L.first[0], L.count[0], A, C, in_rp, lo);
            for (int c = 1; c <= 2; ++c)
                if (L.count[c])
                    hipLaunchKernelGGL(k_cdlp_first_wave, dim3(grid_cap(ceil_div(L.count[c], kBT / 64))), dim3(kBT), 0,
                                       s, L.first[c], L.count[c], A, C, in_rp, lo);
            for (int c = 3; c <= 4; ++c)
                if (L.count[c])
                    hipLaunchKernelGGL(k_cdlp_first_block, dim3(grid_cap(L.count[c])), dim3(kBT), 0, s, L.first[c],
                                       L.count[c], A, C, in_rp, lo);
            if (timing) GHET_HIP(hipEventRecord(kev[2 * it + 1], s));
            continue;
        }
        if (!ranked) {
            hipLaunchKernelGGL(k_cdlp_to_rank, dim3(gn), dim3(256), 0, s, L.perm, lo, n, la, lb);
            ranked = true;
        }
        const int32_t* A = L.rowptr_r;
        const int32_t* C = L.col_r;
        if (L.count[0])
            hipLaunchKernelGGL(k_cdlp_small, dim3(grid_cap(ceil_div(L.count[0], kBT))), dim3(kBT), 0, s, L.list[0],
                               L.count[0], A, C, la, lb);
        if (L.count[1])
            hipLaunchKernelGGL(k_cdlp_wave, dim3(grid_cap(ceil_div(L.count[1], kBT / 64))), dim3(kBT), 0, s,
                               L.list[1], L.count[1], A, C, la, lb);
        if (L.count[2])
            hipLaunchKernelGGL((k_cdlp_lds<kBT, kLdsSlots>), dim3(grid_cap(L.count[2])), dim3(kBT), 0, s, L.list[2],
                               L.count[2], A, C, la, lb);
        if (L.count[3])
            hipLaunchKernelGGL((k_cdlp_lds<kBT, kLdsSlots>), dim3(grid_cap(L.count[3])), dim3(kBT), 0, s, L.list[3],
                               L.count[3], A, C, la, lb);
        if (L.count[4]) {
            hipLaunchKernelGGL(k_cdlp_hub_count, dim3(L.hub_chunks), dim3(kBigBT), 0, s, L.list[4], L.chunk_hub,
                               L.chunk_beg, A, C, la, L.pairs, L.chunk_pairs, L.hub_pairs);
            hipLaunchKernelGGL(k_cdlp_hub_plan, dim3(1), dim3(1024), 0, s, L.hub_pairs, L.count[4], L.item_off);
            hipLaunchKernelGGL(k_cdlp_hub_merge, dim3(merge_blocks), dim3(kBigBT), 0, s, L.item_off, L.count[4],
                               L.hub_chunk0, L.pairs, L.chunk_pairs, merge_cap, L.hub_best);
            hipLaunchKernelGGL(k_cdlp_hub_final, dim3(grid_cap(ceil_div(L.count[4], 256))), dim3(256), 0, s, L.list[4],
                               L.count[4], L.hub_best, lb);
        }
        if (timing) GHET_HIP(hipEventRecord(kev[2 * it + 1], s));
        std::swap(la, lb);
    }
    if (ranked) hipLaunchKernelGGL(k_cdlp_from_rank, dim3(gn), dim3(256), 0, s, L.perm, la, n, lo);
    GHET_HIP(hipEventRecord(g->ev_end, s));
    GHET_HIP(hipGetLastError());
    if (label_out) GHET_CHECK(copy_out(g, label_out, lo, 4 * size_t(n), s));
    GHET_HIP(hipStreamSynchronize(s));
    if (st) {
        double hot = 0;
        for (int32_t it = 0; it < iters; ++it) {
            float kms = 0;
            GHET_HIP(hipEventElapsedTime(&kms, kev[2 * it], kev[2 * it + 1]));
            hot += kms;
        }
        float ms = 0;
        GHET_HIP(hipEventElapsedTime(&ms, g->ev_begin, g->ev_end));
        st->device_ms = ms;
        st->hot_kernel_ms = hot;
        st->hot_kernel_launches = iters;
        st->levels = iters;
        for (int c = 0; c < 3; ++c) st->extra[c] = L.count[c];
        st->extra[3] = L.count[3] + L.count[4];  // vertices counted by blocks of 1024 threads (16K tables, hubs)
    }
    return GHET_OK;
}

}  // namespace ghet
