// Breadth-first search on gfx950: direction-optimising, three frontier modes.
//
// Reference: csr_bfs / Cuda_BFS (GPU_BFS_csr.cuh:258-335) expands a queue into
// an edge queue (gather) and filters it back into a vertex queue (contract),
// two launches and four blocking copies per level, top-down only.  Semantics
// are those of CPU_BFS (CPU_BFS.hpp:9-41): depth 0 at the source, INT_MAX when
// unreached.  BFS depths are unique, so every mode below gives bit-identical
// output; the mode only changes speed.
//
// Modes per level:
//   PUSH    queue frontier; edge-balanced expansion (2048-edge tiles, LDS-staged
//           owner search and LDS-staged next queue).  Every queue is written
//           with its push entries {first edge, row offset} by its producer,
//           so a push level needs no degree-scan launch.
//   PULL    bitmap frontier; each unvisited vertex scans its in-list and stops at
//           the first frontier parent; one wave owns 64 vertices = one 64-bit
//           word of the bitmaps, so the next frontier is written without atomics.
//   BITMAP  bitmap frontier, top-down: lanes expand light vertices, waves expand
//           medium ones, the whole grid shares hub (deg > 4096) vertices.
#include <cmath>

#include "graph_internal.hpp"
#include "sched.hpp"

namespace ghet {
namespace {

// Per-level counters.  hubs/alloc/conv hand out queue positions (one atomic
// per hub / LDS flush / block); the sums are spread over kSlots slots:
// v[0] vertices discovered, v[1] their out-edges, v[2] their in-edges
// (v[7] and done count block arrivals for the mailbox post).
struct BfsCtr {
    int32_t hubs;
    unsigned int done;         // blocks finished (mailbox ticket)
    unsigned long long alloc;  // next queue: (entries << 32) | edges reserved so far
    unsigned long long conv;   // this level's queue from a bitmap: the same packing
    int32_t pad[10];
    CtrSlot slots[kSlots];
};
static_assert(offsetof(BfsCtr, slots) == 64, "counter header of 64 bytes");
constexpr int kCtrWords = sizeof(BfsCtr) / 8;

constexpr int kBT = 256;
constexpr int kItems = 8;
constexpr int kTile = kBT * kItems;  // edges per push tile
constexpr int kOut = 2 * kTile;      // LDS output queue capacity
static_assert(kOut >= kTile, "a push tile can stage kTile discoveries");
constexpr int kHubDeg = 4096;
constexpr int kMaxBlocks = 2048;     // persistent grid cap (8 blocks per CU)
constexpr int kPullSets = 4;         // vertices a k_pull thread scans side by side (2: 5 % slower at s24, 8: 9 %)

__device__ __forceinline__ int upper_bound_i32(const int32_t* a, int len, int32_t key) {
    int lo = 0, hi = len;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (a[mid] <= key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Entries of a push queue: e.x = the entry's first edge (ascending over the
// queue), e.y = its row offset.  Number of entries with e.x <= key.
__device__ __forceinline__ int upper_bound_ent(const int2* a, int len, int32_t key) {
    int lo = 0, hi = len;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (a[mid].x <= key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Level-0 state in one pass: depth INT_MAX except the source (0), visited =
// the no-in-edge mask (none without an in-CSR) plus the source, queue =
// {src} with its push entry, both counter buffers zero, and clear_bm (if
// given) zero.
__global__ void __launch_bounds__(kBT) k_bfs_init(int32_t n, int64_t words, const uint32_t* __restrict__ skip,
                                                  int32_t src, const int32_t* __restrict__ rowptr,
                                                  int32_t* __restrict__ depth, uint32_t* __restrict__ visited,
                                                  int32_t* __restrict__ queue, int2* __restrict__ ent, BfsCtr* ctrs,
                                                  uint32_t* __restrict__ clear_bm) {
    if (blockIdx.x == 0) {  // both counter buffers
        unsigned long long* w = reinterpret_cast<unsigned long long*>(ctrs);
        for (int i = threadIdx.x; i < 2 * kCtrWords; i += kBT) w[i] = 0;
    }
    const int64_t len = n > words ? n : words;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x; i < len;
         i += static_cast<int64_t>(gridDim.x) * kBT) {
        if (i < n) depth[i] = i == src ? 0 : kIntInf;
        if (i < words) {
            uint32_t w = skip ? skip[i] : 0u;
            if (i == (src >> 5)) w |= 1u << (src & 31);
            visited[i] = w;
            if (clear_bm) clear_bm[i] = 0u;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        queue[0] = src;
        ent[0] = make_int2(0, rowptr[src]);
    }
}

__global__ void k_zero_ctr(BfsCtr* c) {
    unsigned long long* w = reinterpret_cast<unsigned long long*>(c);
    for (int i = threadIdx.x; i < kCtrWords; i += blockDim.x) w[i] = 0;
}

// Block 0 clears the counters the next level will use (the other buffer of
// the pair; the host has already copied it), so that level needs no reset
// launch.
__device__ __forceinline__ void zero_next(BfsCtr* c) {
    if (c && blockIdx.x == 0) {
        unsigned long long* w = reinterpret_cast<unsigned long long*>(c);
        for (int i = threadIdx.x; i < kCtrWords; i += blockDim.x) w[i] = 0;
    }
}

// Block totals of (count, out-edges, in-edges) into the block's slot.
// Returns (in thread 0) a token of the atomics' results for post_level.
__device__ __forceinline__ unsigned long long block_report(unsigned long long cnt, unsigned long long od,
                                                           unsigned long long id, BfsCtr* ctr,
                                                           unsigned long long* s_red) {
    cnt = block_sum_u64<kBT>(cnt, s_red);
    od = block_sum_u64<kBT>(od, s_red);
    id = block_sum_u64<kBT>(id, s_red);
    unsigned long long token = 0;
    if (threadIdx.x == 0) {
        token ^= slot_add_ret(ctr->slots, 0, cnt);
        token ^= slot_add_ret(ctr->slots, 1, od);
        token ^= slot_add_ret(ctr->slots, 2, id);
    }
    return token;
}

// The level's last block posts (discovered, their out-edges, their in-edges)
// to the host mailbox.  Block-uniform; mail == nullptr posts nothing.
__device__ __forceinline__ void post_level(BfsCtr* ctr, HostMail* mail, unsigned long long seq,
                                           unsigned long long token) {
    if (!mail) return;
    if (arrive_last(ctr->slots, &ctr->done, token) && threadIdx.x < 64) {  // wave 0: lane i reads slot i
        static_assert(kSlots == 64, "one slot per lane");
        const unsigned long long a = wave_sum(load_agent(&ctr->slots[threadIdx.x].v[0]));
        const unsigned long long b = wave_sum(load_agent(&ctr->slots[threadIdx.x].v[1]));
        const unsigned long long c = wave_sum(load_agent(&ctr->slots[threadIdx.x].v[2]));
        if (threadIdx.x == 0) {
            mail->v[0] = a;
            mail->v[1] = b;
            mail->v[2] = c;
            mail_post(mail, seq);
        }
    }
}

// Moves the c staged discoveries to the next queue together with their push
// entries.  Thread t takes the consecutive items [t P, t P + P), P =
// ceil(c / kBT); the block scans their out-degrees and reserves its entry
// and edge ranges with one 64-bit atomic on *alloc ((entries << 32) |
// edges), so the next level's entries ascend in both, whatever order the
// blocks flush in.  Adds the items and their out- and in-degrees to the
// block's counts.  Block-uniform; flushes only above limit staged items.
__device__ __forceinline__ void flush_entries(int32_t* buf, int32_t* cnt, int limit, unsigned long long* alloc,
                                              int32_t* __restrict__ q, int2* __restrict__ ent,
                                              const int32_t* __restrict__ rowptr,
                                              const int32_t* __restrict__ in_rowptr, int32_t* s_wave,
                                              unsigned long long* s_base, unsigned long long& items,
                                              unsigned long long& od, unsigned long long& id) {
    __syncthreads();
    const int c = *cnt;
    if (c == 0 || c <= limit) {
        __syncthreads();
        return;
    }
    const int lane = lane_id(), wave = threadIdx.x >> 6;
    const int per = (c + kBT - 1) / kBT;
    const int j0 = threadIdx.x * per, j1 = min(j0 + per, c);
    int32_t sum = 0;
    for (int j = j0; j < j1; j += 4) {  // four rows' offsets in flight
        int32_t rb[4], re[4], ib[4], ie[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int32_t v = buf[min(j + k, j1 - 1)];
            rb[k] = rowptr[v];
            re[k] = rowptr[v + 1];
            ib[k] = in_rowptr ? in_rowptr[v] : 0;
            ie[k] = in_rowptr ? in_rowptr[v + 1] : 0;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (j + k < j1) {
                sum += re[k] - rb[k];
                id += ie[k] - ib[k];
                ++items;
            }
    }
    od += sum;
    int32_t incl = sum;
    for (int off = 1; off < 64; off <<= 1) {
        const int32_t y = __shfl_up(incl, off);
        if (lane >= off) incl += y;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    int32_t run = incl - sum, total = 0;
    for (int w = 0; w < kBT / 64; ++w) {
        if (w < wave) run += s_wave[w];
        total += s_wave[w];
    }
    if (threadIdx.x == 0)
        *s_base = atomicAdd(alloc, (static_cast<unsigned long long>(c) << 32) | static_cast<uint32_t>(total));
    __syncthreads();
    int32_t slot = static_cast<int32_t>(*s_base >> 32) + j0;
    int32_t first = static_cast<int32_t>(*s_base & 0xFFFFFFFFu) + run;
    for (int j = j0; j < j1; j += 4) {
        int32_t v[4], rb[4], re[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            v[k] = buf[min(j + k, j1 - 1)];
            rb[k] = rowptr[v[k]];
            re[k] = rowptr[v[k] + 1];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (j + k < j1) {
                q[slot] = v[k];
                ent[slot] = make_int2(first, rb[k]);
                ++slot;
                first += re[k] - rb[k];
            }
    }
    __syncthreads();
    if (threadIdx.x == 0) *cnt = 0;
    __syncthreads();
}

// PUSH: persistent blocks over tiles of tile_edges (<= kTile) consecutive
// frontier edges; entry j of the level covers edges [ent[j].x, ent[j+1].x)
// (E for the last) of the row starting at ent[j].y.  Discoveries go to the
// next queue with their entries and, when next_bm is given (a cleared
// bitmap), to the next frontier bitmap too, so a pull level after a large
// push level needs no queue-to-bitmap pass.  Discoveries are staged in LDS
// and flushed to the next queue in bulk.
__global__ void __launch_bounds__(kBT) k_push_expand(const int2* __restrict__ ent, int32_t F, int32_t E,
                                                     int32_t tile_edges, const int32_t* __restrict__ rowptr,
                                                     const int32_t* __restrict__ col, uint32_t* visited,
                                                     int32_t* __restrict__ depth, int32_t next_level,
                                                     int32_t* __restrict__ next_q, int2* __restrict__ next_ent,
                                                     BfsCtr* ctr, const int32_t* __restrict__ in_rowptr,
                                                     BfsCtr* next_ctr, uint32_t* next_bm, HostMail* mail,
                                                     unsigned long long seq) {
    __shared__ int32_t s_scan[kTile];
    __shared__ int32_t s_start[kTile];
    __shared__ int32_t s_out[kOut];
    __shared__ int32_t s_i0, s_span, s_cnt;
    __shared__ int32_t s_wave[kBT / 64];
    __shared__ unsigned long long s_base;
    __shared__ unsigned long long s_red[kBT / 64];
    const int tid = threadIdx.x;
    if (tid == 0) s_cnt = 0;
    zero_next(next_ctr);
    __syncthreads();  // counter zeroed before any wave appends
    unsigned long long cnt = 0, od = 0, id = 0;
    for (int64_t tile = blockIdx.x; tile * tile_edges < E; tile += gridDim.x) {
        const int32_t t0 = static_cast<int32_t>(tile * tile_edges);
        const int32_t t1 = static_cast<int32_t>(std::min<int64_t>(static_cast<int64_t>(t0) + tile_edges, E));
        if (tid == 0) {
            // ent[0].x = 0 <= t0 < t1 <= E, so both searches land in [0, F)
            int i0 = upper_bound_ent(ent, F, t0) - 1;
            int i1 = upper_bound_ent(ent, F, t1 - 1) - 1;
            s_i0 = i0;
            s_span = i1 - i0 + 1;
        }
        __syncthreads();
        const int i0 = s_i0, span = s_span;
        const bool use_lds = span <= kTile;
        if (use_lds) {
            for (int j = tid; j < span; j += kBT) {
                const int2 en = ent[i0 + j];
                s_scan[j] = en.x;
                s_start[j] = en.y;
            }
        }
        __syncthreads();
        // Each stage runs over the thread's kItems edges together, so the
        // target loads, then the visited-word loads, are in flight at once.
        int32_t pos[kItems], v[kItems];
        bool ok[kItems];
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            const int32_t e = t0 + k * kBT + tid;
            ok[k] = e < t1;
            pos[k] = 0;  // E > 0, so edge 0 exists: masked lanes load it
            if (ok[k]) {
                if (use_lds) {
                    const int j = upper_bound_i32(s_scan, span, e) - 1;
                    pos[k] = s_start[j] + (e - s_scan[j]);
                } else {
                    const int j = upper_bound_ent(ent + i0, span, e) - 1;
                    const int2 en = ent[i0 + j];
                    pos[k] = en.y + (e - en.x);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < kItems; ++k) v[k] = col[pos[k]];
        uint32_t vis[kItems];
#pragma unroll
        for (int k = 0; k < kItems; ++k) vis[k] = ok[k] ? visited[v[k] >> 5] : ~0u;
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            const uint32_t bit = 1u << (v[k] & 31);
            bool disc = false;
            if (!(vis[k] & bit)) {
                const uint32_t old = atomicOr(&visited[v[k] >> 5], bit);
                if (!(old & bit)) {
                    depth[v[k]] = next_level;
                    disc = true;
                    if (next_bm) atomicOr(&next_bm[v[k] >> 5], bit);
                }
            }
            lq_push(s_out, &s_cnt, disc, v[k]);
        }
        flush_entries(s_out, &s_cnt, kOut - kTile, &ctr->alloc, next_q, next_ent, rowptr, in_rowptr, s_wave, &s_base,
                      cnt, od, id);
    }
    flush_entries(s_out, &s_cnt, 0, &ctr->alloc, next_q, next_ent, rowptr, in_rowptr, s_wave, &s_base, cnt, od, id);
    post_level(ctr, mail, seq, block_report(cnt, od, id, ctr, s_red));
}

// Bits of the vertices without in-edges (and of the padding past n): no BFS
// level can reach them, so the visited bitmap starts with them set and pull
// levels skip them.  Depths are untouched (they stay unreached unless one is
// the source), and the level statistics are counted on the host.
__global__ void __launch_bounds__(kBT) k_skip_mask(int32_t n, const int32_t* __restrict__ in_rowptr, int64_t words,
                                                   uint32_t* __restrict__ mask) {
    for (int64_t wd = static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x; wd < words;
         wd += static_cast<int64_t>(gridDim.x) * kBT) {
        uint32_t m = 0;
        for (int b = 0; b < 32; ++b) {
            const int64_t v = wd * 32 + b;
            if (v >= n || in_rowptr[v + 1] == in_rowptr[v]) m |= 1u << b;
        }
        mask[wd] = m;
    }
}

// First in-neighbour of every vertex (-1 without in-edges): the first probe
// of a pull scan then reads a coalesced array instead of one scattered in-CSR
// line per vertex.
__global__ void __launch_bounds__(kBT) k_first_in(int32_t n, const int32_t* __restrict__ in_rowptr,
                                                  const int32_t* __restrict__ in_col, int32_t* __restrict__ first) {
    for (int64_t v = static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x; v < n;
         v += static_cast<int64_t>(gridDim.x) * kBT) {
        const int32_t b = in_rowptr[v];
        first[v] = in_rowptr[v + 1] > b ? in_col[b] : -1;
    }
}

// PULL: bottom-up over the in-CSR.  Each block iteration covers kPullSets
// sets of 256 consecutive vertices; in each set wave w owns 64 of them =
// bitmap words 2g, 2g+1, so the next frontier and the visited bits are
// written without atomics.  A thread scans the in-lists of its kPullSets
// vertices side by side, one in-edge of each per step, so their loads are in
// flight together without probing more edges than a one-vertex scan.
template <int kPullSets>
__global__ void __launch_bounds__(kBT) k_pull(int32_t n, const int32_t* __restrict__ in_rowptr,
                                              const int32_t* __restrict__ in_col, const int32_t* __restrict__ first_in,
                                              const uint32_t* __restrict__ cur_bm,
                                              uint32_t* __restrict__ next_bm, uint32_t* __restrict__ visited,
                                              int32_t* __restrict__ depth, int32_t next_level, BfsCtr* ctr,
                                              const int32_t* __restrict__ out_rowptr, BfsCtr* next_ctr,
                                              HostMail* mail, unsigned long long seq) {
    __shared__ unsigned long long s_red[kBT / 64];
    const int lane = lane_id();
    zero_next(next_ctr);
    unsigned long long od = 0, id = 0, cnt = 0;
    for (int64_t base = static_cast<int64_t>(blockIdx.x) * kBT * kPullSets; base < n;
         base += static_cast<int64_t>(gridDim.x) * kBT * kPullSets) {
        int64_t v[kPullSets];
        uint32_t vis[kPullSets];
        int32_t i[kPullSets], e[kPullSets];
        bool found[kPullSets];
#pragma unroll
        for (int p = 0; p < kPullSets; ++p) {
            v[p] = base + p * kBT + threadIdx.x;
            vis[p] = visited[2 * (v[p] >> 6) + (lane >> 5)];
        }
        int32_t fst[kPullSets], ob[kPullSets], oe[kPullSets];
#pragma unroll
        for (int p = 0; p < kPullSets; ++p) {
            const bool need = v[p] < n && !((vis[p] >> (lane & 31)) & 1u);
            i[p] = need ? in_rowptr[v[p]] : 0;
            e[p] = need ? in_rowptr[v[p] + 1] : 0;
            fst[p] = need ? first_in[v[p]] : -1;
            // the out-offsets load now, beside the in-offsets, instead of
            // after the scan: a found vertex's degrees add no round trip
            ob[p] = need ? out_rowptr[v[p]] : 0;
            oe[p] = need ? out_rowptr[v[p] + 1] : 0;
            found[p] = false;
        }
        int32_t indeg[kPullSets];
#pragma unroll
        for (int p = 0; p < kPullSets; ++p) indeg[p] = e[p] - i[p];
        {  // first probe: the coalesced first in-neighbour
            uint32_t wd[kPullSets];
#pragma unroll
            for (int p = 0; p < kPullSets; ++p) wd[p] = fst[p] >= 0 ? cur_bm[fst[p] >> 5] : 0u;
#pragma unroll
            for (int p = 0; p < kPullSets; ++p) {
                if (fst[p] >= 0) {
                    found[p] = ((wd[p] >> (fst[p] & 31)) & 1u) != 0;
                    i[p] = found[p] ? e[p] : i[p] + 1;
                }
            }
        }
        for (bool more = true; more;) {
            int32_t u[kPullSets];
            uint32_t wd[kPullSets];
#pragma unroll
            for (int p = 0; p < kPullSets; ++p) u[p] = i[p] < e[p] ? in_col[i[p]] : -1;
#pragma unroll
            for (int p = 0; p < kPullSets; ++p) wd[p] = u[p] >= 0 ? cur_bm[u[p] >> 5] : 0u;
            more = false;
#pragma unroll
            for (int p = 0; p < kPullSets; ++p) {
                if (u[p] >= 0) {
                    found[p] = ((wd[p] >> (u[p] & 31)) & 1u) != 0;
                    i[p] = found[p] ? e[p] : i[p] + 1;  // a found vertex stops scanning
                    more |= i[p] < e[p];
                }
            }
        }
#pragma unroll
        for (int p = 0; p < kPullSets; ++p) {
            const int64_t g = v[p] >> 6;
            const uint64_t mask = __ballot(found[p]);
            if (lane == 0) {
                next_bm[2 * g] = static_cast<uint32_t>(mask);
                visited[2 * g] = vis[p] | static_cast<uint32_t>(mask);
            } else if (lane == 32) {
                next_bm[2 * g + 1] = static_cast<uint32_t>(mask >> 32);
                visited[2 * g + 1] = vis[p] | static_cast<uint32_t>(mask >> 32);
            }
            if (found[p]) {
                depth[v[p]] = next_level;
                ++cnt;
                od += oe[p] - ob[p];
                id += indeg[p];
            }
        }
    }
    post_level(ctr, mail, seq, block_report(cnt, od, id, ctr, s_red));
}

__device__ __forceinline__ bool visit(int32_t v, uint32_t* visited, uint32_t* next_bm, int32_t* depth,
                                      int32_t next_level) {
    const uint32_t bit = 1u << (v & 31);
    if (visited[v >> 5] & bit) return false;
    uint32_t old = atomicOr(&visited[v >> 5], bit);
    if (old & bit) return false;
    depth[v] = next_level;
    atomicOr(&next_bm[v >> 5], bit);
    return true;
}

// BITMAP: top-down from a frontier bitmap.  Light vertices (deg <= 32) are
// expanded by their own lane, medium ones by the wave, hubs are queued for
// k_hub_expand.
__global__ void __launch_bounds__(kBT) k_bitmap_topdown(int32_t n, const int32_t* __restrict__ rowptr,
                                                        const int32_t* __restrict__ col,
                                                        const uint32_t* __restrict__ cur_bm, uint32_t* next_bm,
                                                        uint32_t* visited, int32_t* __restrict__ depth,
                                                        int32_t next_level, BfsCtr* ctr, int32_t* hub_q,
                                                        const int32_t* __restrict__ in_rowptr) {
    __shared__ unsigned long long s_red[kBT / 64];
    const int lane = lane_id();
    unsigned long long od = 0, id = 0, cnt = 0;
    auto count = [&](int32_t v) {
        ++cnt;
        od += rowptr[v + 1] - rowptr[v];
        if (in_rowptr) id += in_rowptr[v + 1] - in_rowptr[v];
    };
    for (int64_t base = static_cast<int64_t>(blockIdx.x) * kBT; base < n; base += static_cast<int64_t>(gridDim.x) * kBT) {
        const int64_t u = base + threadIdx.x;
        const uint32_t word = cur_bm[2 * (u >> 6) + (lane >> 5)];
        const bool in_f = (word >> (lane & 31)) & 1u;
        if (__ballot(in_f) == 0) continue;
        int32_t b = 0, e = 0;
        if (in_f) {
            b = rowptr[u];
            e = rowptr[u + 1];
        }
        const int32_t deg = e - b;
        const bool light = in_f && deg <= 32;
        const bool hub = in_f && deg > kHubDeg;
        if (light)
            for (int32_t i = b; i < e; ++i) {
                int32_t v = col[i];
                if (visit(v, visited, next_bm, depth, next_level)) count(v);
            }
        if (hub) hub_q[atomicAdd(&ctr->hubs, 1)] = static_cast<int32_t>(u);
        uint64_t med = __ballot(in_f && !light && !hub);
        while (med) {
            const int src_lane = __ffsll(static_cast<unsigned long long>(med)) - 1;
            med &= med - 1;
            const int32_t mb = __shfl(b, src_lane), me = __shfl(e, src_lane);
            for (int32_t i = mb + lane; i < me; i += 64) {
                int32_t v = col[i];
                if (visit(v, visited, next_bm, depth, next_level)) count(v);
            }
        }
    }
    block_report(cnt, od, id, ctr, s_red);
}

// Hub vertices queued by k_bitmap_topdown: every block takes a slice of every hub.
__global__ void __launch_bounds__(kBT) k_hub_expand(const int32_t* __restrict__ rowptr,
                                                    const int32_t* __restrict__ col, uint32_t* next_bm,
                                                    uint32_t* visited, int32_t* __restrict__ depth,
                                                    int32_t next_level, BfsCtr* ctr, const int32_t* hub_q,
                                                    const int32_t* __restrict__ in_rowptr, HostMail* mail,
                                                    unsigned long long seq) {
    __shared__ unsigned long long s_red[kBT / 64];
    const int32_t H = ctr->hubs;
    unsigned long long od = 0, id = 0, cnt = 0;
    for (int32_t h = 0; h < H; ++h) {
        const int32_t u = hub_q[h];
        const int32_t b = rowptr[u], e = rowptr[u + 1];
        for (int64_t i = b + static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x; i < e;
             i += static_cast<int64_t>(gridDim.x) * kBT) {
            int32_t v = col[i];
            if (visit(v, visited, next_bm, depth, next_level)) {
                ++cnt;
                od += rowptr[v + 1] - rowptr[v];
                if (in_rowptr) id += in_rowptr[v + 1] - in_rowptr[v];
            }
        }
    }
    post_level(ctr, mail, seq, block_report(cnt, od, id, ctr, s_red));
}

__global__ void k_queue_to_bitmap(const int32_t* __restrict__ q, int32_t F, uint32_t* bm) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < F) {
        int32_t v = q[i];
        atomicOr(&bm[v >> 5], 1u << (v & 31));
    }
}

// Frontier bitmap -> queue with its push entries: each thread takes 4 bitmap
// words (128 vertices), the block scans their popcounts and out-degrees and
// reserves its entry and edge ranges with one 64-bit atomic ((entries << 32)
// | edges), so the entries ascend in both.  One pass, no loop: words is a
// multiple of 8, so a thread's 4 words are either all present or all past
// the end.
__global__ void __launch_bounds__(kBT) k_bitmap_to_queue(int64_t words, int32_t n, const uint32_t* __restrict__ bm,
                                                         const int32_t* __restrict__ rowptr, int32_t* __restrict__ q,
                                                         int2* __restrict__ ent, unsigned long long* counter) {
    __shared__ int32_t s_wave[kBT / 64], s_wdeg[kBT / 64];
    __shared__ unsigned long long s_base;
    const int lane = lane_id(), wave = threadIdx.x >> 6;
    const int64_t w0 = (static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x) * 4;
    uint4 x = make_uint4(0u, 0u, 0u, 0u);
    if (w0 < words) x = *reinterpret_cast<const uint4*>(bm + w0);
    // bits of ids >= n (padding) never count
    auto valid = [n](int64_t wi) -> uint32_t {
        const int64_t lo = wi * 32;
        return lo + 32 <= n ? ~0u : (lo >= n ? 0u : (1u << (n - lo)) - 1u);
    };
    x.x &= valid(w0);
    x.y &= valid(w0 + 1);
    x.z &= valid(w0 + 2);
    x.w &= valid(w0 + 3);
    const int c = __popc(x.x) + __popc(x.y) + __popc(x.z) + __popc(x.w);
    const uint32_t ws[4] = {x.x, x.y, x.z, x.w};
    int32_t dsum = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        for (uint32_t b = ws[k]; b; b &= b - 1) {
            const int64_t v = (w0 + k) * 32 + (__ffs(b) - 1);
            dsum += rowptr[v + 1] - rowptr[v];
        }
    int incl = c, dincl = dsum;
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(incl, off), yd = __shfl_up(dincl, off);
        if (lane >= off) {
            incl += y;
            dincl += yd;
        }
    }
    if (lane == 63) {
        s_wave[wave] = incl;
        s_wdeg[wave] = dincl;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0, td = 0;
        for (int w = 0; w < kBT / 64; ++w) {
            t += s_wave[w];
            td += s_wdeg[w];
        }
        s_base = t > 0 ? atomicAdd(counter, (static_cast<unsigned long long>(t) << 32) | static_cast<uint32_t>(td)) : 0;
    }
    __syncthreads();
    int off = static_cast<int32_t>(s_base >> 32) + incl - c;
    int32_t first = static_cast<int32_t>(s_base & 0xFFFFFFFFu) + dincl - dsum;
    for (int w = 0; w < wave; ++w) {
        off += s_wave[w];
        first += s_wdeg[w];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        for (uint32_t b = ws[k]; b; b &= b - 1) {
            const int32_t v = static_cast<int32_t>((w0 + k) * 32 + (__ffs(b) - 1));
            const int32_t rb = rowptr[v];
            q[off] = v;
            ent[off] = make_int2(first, rb);
            ++off;
            first += rowptr[v + 1] - rb;
        }
    }
}

// Parent = smallest u with an edge u->v and depth[u] == depth[v] - 1
// (atomicMin over out-edges: deterministic).  Lanes expand light rows,
// the wave expands heavy ones.  pre must hold INT_MAX beforehand.
__global__ void __launch_bounds__(kBT) k_bfs_parent(int32_t n, const int32_t* __restrict__ rowptr,
                                                    const int32_t* __restrict__ col,
                                                    const int32_t* __restrict__ depth, int32_t* pre) {
    const int lane = lane_id();
    const int64_t groups = ceil_div(n, 64);
    const int64_t wave_global = (static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x) >> 6;
    const int64_t wave_stride = (static_cast<int64_t>(gridDim.x) * kBT) >> 6;
    for (int64_t g = wave_global; g < groups; g += wave_stride) {
        const int64_t u = g * 64 + lane;
        int32_t b = 0, e = 0, du = kIntInf;
        if (u < n) {
            du = depth[u];
            if (du != kIntInf) {
                b = rowptr[u];
                e = rowptr[u + 1];
            }
        }
        const bool light = (e - b) <= 32;
        if (light)
            for (int32_t i = b; i < e; ++i)
                if (depth[col[i]] == du + 1) atomicMin(&pre[col[i]], static_cast<int32_t>(u));
        uint64_t heavy = __ballot(!light);
        while (heavy) {
            const int src = __ffsll(static_cast<unsigned long long>(heavy)) - 1;
            heavy &= heavy - 1;
            const int32_t hb = __shfl(b, src), he = __shfl(e, src), hd = __shfl(du, src);
            const int32_t hu = static_cast<int32_t>(g * 64 + src);
            for (int32_t i = hb + lane; i < he; i += 64)
                if (depth[col[i]] == hd + 1) atomicMin(&pre[col[i]], hu);
        }
    }
}

__global__ void k_pre_finish(int32_t* pre, int32_t n) {
    for (int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < n;
         v += static_cast<int64_t>(gridDim.x) * blockDim.x)
        if (pre[v] == kIntInf) pre[v] = -1;
}

unsigned grid_for_vertices(int64_t n) {
    return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, kBT), 4096)));
}

}  // namespace

int run_bfs(ghet_graph* g, int32_t src, int32_t* depth_out, int32_t* pre_out, const ghet_opts* o,
            ghet_run_stats* st) {
    const int32_t n = g->n;
    if (src < 0 || src >= n) return set_error(GHET_ERR_INVALID, "source %d out of range [0,%d)", src, n);
    if (!g->out.present()) return set_error(GHET_ERR_INVALID, "graph has no out-CSR");
    hipStream_t s = pick_stream(g, o);
    int schedule = (o && o->size >= sizeof(ghet_opts)) ? o->schedule : GHET_SCHED_HEURISTIC;
    const bool have_in = g->in.present();
    if (!have_in && (schedule == GHET_SCHED_PULL || schedule == GHET_SCHED_REFERENCE)) schedule = GHET_SCHED_PUSH;

    // 8 words per 256 vertices, padded to whole k_pull block iterations
    // (kPullSets * 256 vertices), so no iteration indexes past the end
    const int64_t words = 8 * kPullSets * ceil_div(std::max(n, 1), 256 * kPullSets);
    int32_t *depth, *qa, *qb, *hubq;
    int2 *enta, *entb;  // push entries of qa / qb
    uint32_t *visited, *bma, *bmb;
    BfsCtr *ctrs, *ctr;
    void* p;
    GHET_CHECK(g->ws.get(WS_DEPTH, 4 * size_t(n), &p)); depth = static_cast<int32_t*>(p);
    GHET_CHECK(g->ws.get(WS_VISITED, 4 * size_t(words), &p)); visited = static_cast<uint32_t*>(p);
    GHET_CHECK(g->ws.get(WS_FRONT_A, 4 * size_t(words), &p)); bma = static_cast<uint32_t*>(p);
    GHET_CHECK(g->ws.get(WS_FRONT_B, 4 * size_t(words), &p)); bmb = static_cast<uint32_t*>(p);
    GHET_CHECK(g->ws.get(WS_QUEUE_A, 4 * size_t(n), &p)); qa = static_cast<int32_t*>(p);
    GHET_CHECK(g->ws.get(WS_QUEUE_B, 4 * size_t(n), &p)); qb = static_cast<int32_t*>(p);
    GHET_CHECK(g->ws.get(WS_SCAN, 8 * size_t(std::max(n, 1)), &p)); enta = static_cast<int2*>(p);
    GHET_CHECK(g->ws.get(WS_BFS_START, 8 * size_t(std::max(n, 1)), &p)); entb = static_cast<int2*>(p);
    GHET_CHECK(g->ws.get(WS_MISC, 4 * size_t(n) + 256, &p)); hubq = static_cast<int32_t*>(p);
    // two counter buffers: level L counts into ctrs[L & 1], and its main
    // kernel clears the other one for level L + 1
    GHET_CHECK(g->ws.get(WS_COUNTERS, 2 * sizeof(BfsCtr), &p)); ctrs = static_cast<BfsCtr*>(p);
    ctr = ctrs;

    SchedState sched;
    sched_begin(&sched, schedule, GHET_NET_BFS, n, g->out.m, have_in ? g->in.m : g->out.m);
    ghet_level_trace* trace = (o && o->size >= sizeof(ghet_opts)) ? o->trace : nullptr;
    const int trace_cap = trace ? o->trace_cap : 0;
    int trace_len = 0;
    GHET_CHECK(mail_init(g));

    if (have_in && !g->bfs_first) {
        GHET_HIP(hipMalloc(reinterpret_cast<void**>(&g->bfs_first), 4 * size_t(std::max(n, 1))));
        hipLaunchKernelGGL(k_first_in, dim3(static_cast<unsigned>(std::min<int64_t>(ceil_div(n, kBT), 8192))),
                           dim3(kBT), 0, s, n, g->in.rowptr, g->in.col, g->bfs_first);
        GHET_HIP(hipGetLastError());
    }
    if (have_in && !g->bfs_skip) {
        GHET_HIP(hipMalloc(reinterpret_cast<void**>(&g->bfs_skip), 4 * size_t(words)));
        hipLaunchKernelGGL(k_skip_mask, dim3(static_cast<unsigned>(std::min<int64_t>(ceil_div(words, kBT), 4096))),
                           dim3(kBT), 0, s, n, g->in.rowptr, words, g->bfs_skip);
        GHET_HIP(hipGetLastError());
    }

    GHET_HIP(hipEventRecord(g->ev_begin, s));
    hipLaunchKernelGGL(k_bfs_init, dim3(static_cast<unsigned>(std::min<int64_t>(ceil_div(std::max<int64_t>(n, words), kBT), 8192))),
                       dim3(kBT), 0, s, n, words, have_in ? g->bfs_skip : nullptr, src, g->out.rowptr, depth, visited,
                       qa, enta, ctrs, have_in ? bma : nullptr);
    bool bma_clear = have_in;  // bma is all zero (until a level writes it)
    bool zeroed[2] = {true, true};  // ctrs[i] known to be all zero

    int32_t F = 1;
    int64_t m_f = g->h_out_rowptr[src + 1] - g->h_out_rowptr[src];
    int64_t visited_in = have_in ? (g->h_in_rowptr[src + 1] - g->h_in_rowptr[src]) : 0;
    int64_t n_visited = 1;
    bool have_q = true, have_bm = false;
    int32_t level = 0;
    int prev_mode = GHET_MODE_PUSH;
    int64_t prev_F = 0;
    const int32_t* in_rp = have_in ? g->in.rowptr : nullptr;
    while (F > 0) {
        LevelInfo li;
        li.n = n;
        li.m = g->out.m;
        li.m_in = have_in ? g->in.m : g->out.m;
        li.level = level;
        li.frontier = F;
        li.frontier_edges = m_f;
        li.unvisited = n - n_visited;
        li.unvisited_in_edges = li.m_in - visited_in;
        li.prev_frontier = prev_F;
        li.prev_mode = prev_mode;
        li.have_in = have_in;
        int mode = sched_pick_bfs(&sched, li);
        const bool tr = trace && trace_len < trace_cap;
        if (tr) {
            sched_features(li, trace[trace_len].features);
            trace[trace_len].threshold = 0;
            GHET_HIP(hipEventRecord(g->ev_k0, s));
        }
        const int32_t next_level = level + 1;
        ctr = ctrs + (level & 1);
        BfsCtr* next_ctr = ctrs + ((level + 1) & 1);
        bool next_zeroed = false;
        unsigned long long seq = 0;  // mailbox sequence of the level's last launch (0: none)
        auto clear = [&]() {
            if (!zeroed[level & 1]) hipLaunchKernelGGL(k_zero_ctr, dim3(1), dim3(512), 0, s, ctr);
            zeroed[level & 1] = true;
        };
        if (mode == GHET_MODE_PUSH && !have_q) {
            clear();
            // counts in ctr->conv: alloc and the slots stay zero
            hipLaunchKernelGGL(k_bitmap_to_queue, dim3(static_cast<unsigned>(ceil_div(words, 4 * kBT))), dim3(kBT), 0, s,
                               words, n, bma, g->out.rowptr, qa, enta, &ctr->conv);
            have_q = true;
        }
        if (mode == GHET_MODE_PUSH) {
            clear();
            const int32_t E = static_cast<int32_t>(m_f);
            // a level this large is usually followed by a pull level: write
            // its frontier bitmap on the way
            const bool want_bm = E > 0 && E >= n / 128 && have_in;
            if (want_bm && !bma_clear) GHET_CHECK(fill<uint32_t>(bma, 0u, words, s));
            if (E > 0) {
                // about 512 tiles while the level is small, kTile-edge tiles
                // over a persistent grid once it is large
                const int32_t tile = static_cast<int32_t>(
                    std::min<int64_t>(kTile, std::max<int64_t>(64, ceil_div(ceil_div(E, 512), 64) * 64)));
                const unsigned grid = static_cast<unsigned>(std::min<int64_t>(ceil_div(E, tile), kMaxBlocks));
                seq = mail_next(g);
                hipLaunchKernelGGL(k_push_expand, dim3(grid), dim3(kBT), 0, s, enta, F, E, tile, g->out.rowptr,
                                   g->out.col, visited, depth, next_level, qb, entb, ctr, in_rp, next_ctr,
                                   want_bm ? bma : nullptr, g->mail_dev, seq);
                next_zeroed = true;
            }
            std::swap(qa, qb);
            std::swap(enta, entb);
            have_q = true;
            have_bm = want_bm;  // bma then holds the same frontier
        } else {
            clear();
            if (!have_bm) {
                GHET_CHECK(fill<uint32_t>(bma, 0u, words, s));
                hipLaunchKernelGGL(k_queue_to_bitmap, dim3(static_cast<unsigned>(ceil_div(F, 256))), dim3(256), 0, s,
                                   qa, F, bma);
                have_bm = true;
            }
            if (mode == GHET_MODE_PULL) {
                seq = mail_next(g);
                hipLaunchKernelGGL(k_pull<kPullSets>, dim3(grid_for_vertices(n)), dim3(kBT), 0, s, n, g->in.rowptr, g->in.col,
                                   g->bfs_first, bma, bmb, visited, depth, next_level, ctr, g->out.rowptr, next_ctr,
                                   g->mail_dev, seq);
                next_zeroed = true;
            } else {
                GHET_CHECK(fill<uint32_t>(bmb, 0u, words, s));
                hipLaunchKernelGGL(k_bitmap_topdown, dim3(grid_for_vertices(n)), dim3(kBT), 0, s, n, g->out.rowptr,
                                   g->out.col, bma, bmb, visited, depth, next_level, ctr, hubq, in_rp);
                seq = mail_next(g);
                hipLaunchKernelGGL(k_hub_expand, dim3(1024), dim3(kBT), 0, s, g->out.rowptr, g->out.col, bmb, visited,
                                   depth, next_level, ctr, hubq, in_rp, g->mail_dev, seq);
            }
            std::swap(bma, bmb);
            have_bm = true;
            have_q = false;
        }
        zeroed[level & 1] = false;
        zeroed[(level + 1) & 1] = next_zeroed;
        bma_clear = false;
        GHET_HIP(hipGetLastError());
        if (tr) GHET_HIP(hipEventRecord(g->ev_k1, s));
        // the level's totals arrive in the mailbox; a push level with no
        // edges discovers nothing
        unsigned long long found = 0, found_out = 0, found_in = 0;
        if (seq) {
            GHET_CHECK(mail_wait(g, s, seq));
            found = g->mail->v[0];
            found_out = g->mail->v[1];
            found_in = g->mail->v[2];
        }
        if (tr) {
            GHET_HIP(hipEventSynchronize(g->ev_k1));
            float ms = 0;
            GHET_HIP(hipEventElapsedTime(&ms, g->ev_k0, g->ev_k1));
            trace[trace_len].mode = mode;
            trace[trace_len].ms = ms;
            ++trace_len;
        }
        prev_F = F;
        prev_mode = mode;
        F = static_cast<int32_t>(found);
        m_f = static_cast<int64_t>(found_out);
        visited_in += static_cast<int64_t>(found_in);
        n_visited += F;
        ++level;
    }
    if (pre_out) {
        GHET_CHECK(g->ws.get(WS_PRE, 4 * size_t(n), &p));
        int32_t* pre = static_cast<int32_t*>(p);
        GHET_CHECK(fill<int32_t>(pre, kIntInf, n, s));
        hipLaunchKernelGGL(k_bfs_parent, dim3(grid_for_vertices(n)), dim3(kBT), 0, s, n, g->out.rowptr, g->out.col,
                           depth, pre);
        hipLaunchKernelGGL(k_pre_finish, dim3(grid_for_vertices(n)), dim3(256), 0, s, pre, n);
    }
    GHET_HIP(hipEventRecord(g->ev_end, s));
    GHET_HIP(hipGetLastError());
    if (depth_out) GHET_CHECK(copy_out(g, depth_out, depth, 4 * size_t(n), s));
    if (pre_out) GHET_CHECK(copy_out(g, pre_out, g->ws.ptr[WS_PRE], 4 * size_t(n), s));
    if (st) {
        CtrSlot* slots = reinterpret_cast<CtrSlot*>(ctr);
        GHET_CHECK(reached_stats_i32(depth, g->out.rowptr, n, slots, s));
        GHET_HIP(hipMemcpyAsync(g->pinned, slots, sizeof(CtrSlot) * kSlots, hipMemcpyDeviceToHost, s));
    }
    GHET_HIP(hipStreamSynchronize(s));
    if (st) {
        float ms = 0;
        GHET_HIP(hipEventElapsedTime(&ms, g->ev_begin, g->ev_end));
        st->device_ms = ms;
        st->levels = level;
        const CtrSlot* hs = reinterpret_cast<const CtrSlot*>(g->pinned);
        st->vertices_reached = static_cast<int64_t>(slot_sum(hs, 0));
        st->edges_traversed = static_cast<int64_t>(slot_sum(hs, 1));
        st->trace_len = trace_len;
        st->status_flags = sched.fallback ? 1 : 0;
    }
    return sched.fallback ? GHET_WARN_FALLBACK : GHET_OK;
}

}  // namespace ghet
