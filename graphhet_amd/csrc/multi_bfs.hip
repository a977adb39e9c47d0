// Batched BFS: up to 256 sources per pass in ONE persistent cooperative
// launch (no host round trip between levels).
//
// No reference counterpart: the reference runs one csr_bfs per source
// (GPU_BFS_csr.cuh:258).  Results are bit-identical to per-source runs.
//
// Bit-parallel traversal: every vertex carries 256-bit sets (4 x u64) --
// seen[v] (sources that reached v) and front[v] (sources whose frontier holds
// v) -- so one pass over the graph advances all 256 searches.  Per level the
// grid agrees on a direction from the frontier's out-edge count:
//   push  frontier vertices OR their set into acc[w] of each out-neighbour
//         (atomicOr), then every vertex takes next = acc & ~seen;
//   pull  every vertex not yet seen by all sources ORs front[u] over its
//         in-neighbours, stopping as soon as nothing unseen is left; in-lists
//         longer than 256 are cut into 2048-edge chunks that any wave takes
//         (OR into acc[v], settled after a grid barrier) -- R-MAT hubs sit at
//         low ids, and one wave walking all of them serialised every pull
//         level on its memory round trips.
// The traversal writes only the 32-byte set of new bits per (level, vertex)
// into a level log (one coalesced store); depths are materialised afterwards
// by a pass that reads a vertex's log column and writes source-major rows
// with coalesced stores.  Levels past the log's depth fall back to direct
// writes.
#include <hip/hip_cooperative_groups.h>

#include "graph_internal.hpp"

namespace cg = cooperative_groups;

namespace ghet {
namespace {

constexpr int kBT = 256;
constexpr int kWords = 4;
constexpr int kPass = 64 * kWords;  // sources per pass
constexpr int kLogLevels = 32;      // levels recorded in the log
constexpr int kPullHeavy = 256;     // pull: longer in-lists are split into chunks
constexpr int kChunkE = 2048;       // in-edges per chunk

struct Bits {
    uint64_t w[kWords];
};

struct MsArgs {
    int32_t n;
    const int32_t *out_rowptr, *out_col, *in_rowptr, *in_col;
    Bits *seen, *fa, *fb, *acc;
    Bits* log;       // [kLogLevels][n]: bits first set at level L+1
    int32_t* out;    // [kp][n] source-major depths (sources and deep levels written directly)
    Bits full;       // active sources of this pass
    int32_t* flags;  // [3] "something new at level L" (L % 3)
    CtrSlot* slots;  // [3][kSlots] per level (v[0] next-frontier out-edges), then [kSlots] totals
                     // (v[1] traversed (source, edge) pairs, v[2] edges scanned by pull, v[3] edges pushed)
    int64_t front_edges0;
    int64_t push_limit;
    int32_t max_levels;
    int32_t* levels_out;
    // optional per-level record (block 0 writes it after each level's
    // barrier): {wall-clock stamp, mode (0 push, 1 pull), frontier out-edges}
    unsigned long long* lvl_rec;
    int32_t lvl_cap;
    const int32_t* heavy;  // [num_heavy] vertices with in-degree > kPullHeavy
    int32_t num_heavy;
    const int2* chunks;    // [num_chunks] {vertex, first in-edge} of their in-lists
    int32_t num_chunks;
};

__device__ __forceinline__ bool any_bits(const Bits& b) { return (b.w[0] | b.w[1] | b.w[2] | b.w[3]) != 0; }

__device__ __forceinline__ Bits wave_or(Bits b) {
    for (int off = 32; off > 0; off >>= 1)
        for (int k = 0; k < kWords; ++k) b.w[k] |= __shfl_xor(b.w[k], off);
    return b;
}

__device__ __forceinline__ bool covers(const Bits& a, const Bits& seen, const Bits& full) {
    bool c = true;
    for (int k = 0; k < kWords; ++k) c &= ((a.w[k] | seen.w[k]) & full.w[k]) == full.w[k];
    return c;
}

// New bits of v at `level`: update seen, log them, count stats.
__device__ __forceinline__ void settle(const MsArgs& a, int64_t v, const Bits& sn, const Bits& nb, int32_t level,
                                       unsigned long long& ne, unsigned long long& tr) {
    Bits s2;
    int pop = 0;
    for (int k = 0; k < kWords; ++k) {
        s2.w[k] = sn.w[k] | nb.w[k];
        pop += __popcll(nb.w[k]);
    }
    a.seen[v] = s2;
    if (level <= kLogLevels) {
        a.log[static_cast<int64_t>(level - 1) * a.n + v] = nb;
    } else {
        for (int k = 0; k < kWords; ++k) {
            uint64_t x = nb.w[k];
            while (x) {
                const int b = __ffsll(static_cast<unsigned long long>(x)) - 1;
                x &= x - 1;
                a.out[static_cast<int64_t>(k * 64 + b) * a.n + v] = level;
            }
        }
    }
    const unsigned long long d = a.out_rowptr[v + 1] - a.out_rowptr[v];
    ne += d;
    tr += d * static_cast<unsigned long long>(pop);
}

__global__ void __launch_bounds__(kBT) k_msbfs(MsArgs a) {
    cg::grid_group grid = cg::this_grid();
    __shared__ unsigned long long s_red[kBT / 64];
    __shared__ int64_t s_front_edges;
    __shared__ int s_stop;
    const int lane = lane_id();
    const int64_t n = a.n;
    CtrSlot* totals = a.slots + 3 * kSlots;
    int64_t front_edges = a.front_edges0;
    int32_t level = 0;
    if (a.lvl_rec && blockIdx.x == 0 && threadIdx.x == 0) a.lvl_rec[0] = wall_clock64();  // start stamp
    for (; level < a.max_levels; ++level) {
        const Bits* front = (level & 1) ? a.fb : a.fa;
        Bits* next = (level & 1) ? a.fa : a.fb;
        Bits* logl = level < kLogLevels ? a.log + static_cast<int64_t>(level) * n : nullptr;
        CtrSlot* slots = a.slots + (level % 3) * kSlots;
        unsigned long long ne = 0, tr = 0, scanned = 0, pushed = 0;
        bool any = false;
        if (front_edges < a.push_limit) {
            // push phase A: OR each frontier set into acc of its out-neighbours
            for (int64_t b0 = static_cast<int64_t>(blockIdx.x) * kBT; b0 < n; b0 += static_cast<int64_t>(gridDim.x) * kBT) {
                const int64_t u = b0 + threadIdx.x;
                Bits f = {};
                int32_t b = 0, e = 0;
                if (u < n) {
                    f = front[u];
                    if (any_bits(f)) {
                        b = a.out_rowptr[u];
                        e = a.out_rowptr[u + 1];
                    }
                }
                pushed += e - b;
                const bool light = e - b <= 32;
                if (light)
                    for (int32_t i = b; i < e; ++i) {
                        Bits* t = a.acc + a.out_col[i];
                        for (int k = 0; k < kWords; ++k)
                            if (f.w[k]) atomicOr(reinterpret_cast<unsigned long long*>(&t->w[k]), f.w[k]);
                    }
                uint64_t heavy = __ballot(!light);
                while (heavy) {
                    const int sl = __ffsll(static_cast<unsigned long long>(heavy)) - 1;
                    heavy &= heavy - 1;
                    const int32_t hb = __shfl(b, sl), he = __shfl(e, sl);
                    Bits hf;
                    for (int k = 0; k < kWords; ++k) hf.w[k] = __shfl(f.w[k], sl);
                    for (int32_t i = hb + lane; i < he; i += 64) {
                        Bits* t = a.acc + a.out_col[i];
                        for (int k = 0; k < kWords; ++k)
                            if (hf.w[k]) atomicOr(reinterpret_cast<unsigned long long*>(&t->w[k]), hf.w[k]);
                    }
                }
            }
            grid.sync();
            // push phase B: next = acc & ~seen, then clear acc
            for (int64_t b0 = static_cast<int64_t>(blockIdx.x) * kBT; b0 < n; b0 += static_cast<int64_t>(gridDim.x) * kBT) {
                const int64_t v = b0 + threadIdx.x;
                if (v >= n) continue;
                const Bits ac = a.acc[v];
                Bits nb = {};
                if (any_bits(ac)) {
                    const Bits sn = a.seen[v];
                    for (int k = 0; k < kWords; ++k) nb.w[k] = ac.w[k] & ~sn.w[k] & a.full.w[k];
                    a.acc[v] = Bits{};
                    if (any_bits(nb)) {
                        settle(a, v, sn, nb, level + 1, ne, tr);
                        any = true;
                    }
                }
                next[v] = nb;
                if (logl && !any_bits(nb)) logl[v] = nb;
            }
        } else {
            // pull: every vertex not yet seen by all sources scans its in-list;
            // in-lists longer than kPullHeavy are left to the chunk pass below
            for (int64_t b0 = static_cast<int64_t>(blockIdx.x) * kBT; b0 < n; b0 += static_cast<int64_t>(gridDim.x) * kBT) {
                const int64_t v = b0 + threadIdx.x;
                Bits sn = {}, acc = {};
                int32_t b = 0, e = 0;
                bool todo = false;
                if (v < n) {
                    sn = a.seen[v];
                    b = a.in_rowptr[v];
                    e = a.in_rowptr[v + 1];
                    todo = e - b <= kPullHeavy && !covers(Bits{}, sn, a.full);
                    if (!todo) e = b;
                }
                const bool light = e - b <= 32;
                if (todo && light)
                    for (int32_t i = b; i < e; i += 4) {
                        // four gathers in flight per lane before the exit test
                        int32_t u[4];
#pragma unroll
                        for (int j = 0; j < 4; ++j) u[j] = i + j < e ? a.in_col[i + j] : -1;
                        Bits f[4];
#pragma unroll
                        for (int j = 0; j < 4; ++j) f[j] = u[j] >= 0 ? front[u[j]] : Bits{};
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                            for (int k = 0; k < kWords; ++k) acc.w[k] |= f[j].w[k];
                        scanned += min(4, e - i);
                        if (covers(acc, sn, a.full)) break;
                    }
                uint64_t heavy = __ballot(!light);
                while (heavy) {
                    const int sl = __ffsll(static_cast<unsigned long long>(heavy)) - 1;
                    heavy &= heavy - 1;
                    const int32_t hb = __shfl(b, sl), he = __shfl(e, sl);
                    Bits hs;
                    for (int k = 0; k < kWords; ++k) hs.w[k] = __shfl(sn.w[k], sl);
                    Bits hacc = {};
                    for (int32_t i0 = hb; i0 < he; i0 += 64) {
                        if (i0 + lane < he) {
                            const Bits f = front[a.in_col[i0 + lane]];
                            ++scanned;
                            for (int k = 0; k < kWords; ++k) hacc.w[k] |= f.w[k];
                        }
                        hacc = wave_or(hacc);
                        if (covers(hacc, hs, a.full)) break;
                    }
                    if (lane == sl) acc = hacc;
                }
                if (v < n) {
                    Bits nb;
                    for (int k = 0; k < kWords; ++k) nb.w[k] = acc.w[k] & ~sn.w[k] & a.full.w[k];
                    if (any_bits(nb)) {
                        settle(a, v, sn, nb, level + 1, ne, tr);
                        any = true;
                    } else if (logl) {
                        logl[v] = nb;
                    }
                    next[v] = nb;
                }
            }
            // Long in-lists: any wave takes any chunk (a hub's list would
            // otherwise serialise on the one wave that owns the hub), ORs the
            // frontier sets of its edges and adds the new bits to acc[v].
            const int64_t wave_id = (static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x) >> 6;
            const int64_t num_waves = (static_cast<int64_t>(gridDim.x) * kBT) >> 6;
            for (int64_t c = wave_id; c < a.num_chunks; c += num_waves) {
                const int2 ch = a.chunks[c];
                const Bits sn = a.seen[ch.x];
                if (covers(Bits{}, sn, a.full)) continue;  // wave-uniform
                const int32_t e1 = min(ch.y + kChunkE, a.in_rowptr[ch.x + 1]);
                Bits acc = {};
                for (int32_t i = ch.y + lane; i < e1; i += 4 * 64) {
                    Bits f[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int32_t ij = i + j * 64;
                        f[j] = ij < e1 ? front[a.in_col[ij]] : Bits{};
                    }
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        for (int k = 0; k < kWords; ++k) acc.w[k] |= f[j].w[k];
                }
                if (lane == 0) scanned += e1 - ch.y;
                acc = wave_or(acc);
                if (lane < kWords) {
                    const uint64_t add = acc.w[lane] & ~sn.w[lane] & a.full.w[lane];
                    if (add) atomicOr(reinterpret_cast<unsigned long long*>(&a.acc[ch.x].w[lane]), add);
                }
            }
            grid.sync();
            // settle the long-list vertices from acc (and clear it for the next level)
            for (int64_t h = static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x; h < a.num_heavy;
                 h += static_cast<int64_t>(gridDim.x) * kBT) {
                const int32_t v = a.heavy[h];
                const Bits ac = a.acc[v];
                if (!any_bits(ac)) continue;
                a.acc[v] = Bits{};
                const Bits sn = a.seen[v];
                Bits nb;
                for (int k = 0; k < kWords; ++k) nb.w[k] = ac.w[k] & ~sn.w[k] & a.full.w[k];
                if (any_bits(nb)) {
                    settle(a, v, sn, nb, level + 1, ne, tr);
                    next[v] = nb;
                    any = true;
                }
            }
        }
        if (__ballot(any) && lane == 0)
            __hip_atomic_store(&a.flags[level % 3], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ne = block_sum_u64<kBT>(ne, s_red);
        tr = block_sum_u64<kBT>(tr, s_red);
        scanned = block_sum_u64<kBT>(scanned, s_red);
        pushed = block_sum_u64<kBT>(pushed, s_red);
        if (threadIdx.x == 0) {
            slot_add(slots, 0, ne);
            slot_add(totals, 1, tr);
            slot_add(totals, 2, scanned);
            slot_add(totals, 3, pushed);
        }
        if (blockIdx.x == 0) {
            // clear the flag and slots of level+1 (last read two levels ago)
            CtrSlot* nxt = a.slots + ((level + 1) % 3) * kSlots;
            for (int i = threadIdx.x; i < kSlots * 8; i += kBT)
                __hip_atomic_store(&nxt[i / 8].v[i % 8], 0ULL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (threadIdx.x == 0)
                __hip_atomic_store(&a.flags[(level + 1) % 3], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        grid.sync();
        if (threadIdx.x == 0) {
            s_stop = __hip_atomic_load(&a.flags[level % 3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
            int64_t fe = 0;
            for (int i = 0; i < kSlots; ++i)
                fe += static_cast<int64_t>(
                    __hip_atomic_load(&slots[i].v[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            s_front_edges = fe;
        }
        __syncthreads();
        if (a.lvl_rec && blockIdx.x == 0 && threadIdx.x == 0 && level + 1 < a.lvl_cap) {
            unsigned long long* r = a.lvl_rec + 3 * static_cast<int64_t>(level + 1);
            r[0] = wall_clock64();
            r[1] = front_edges < a.push_limit ? 0ULL : 1ULL;
            r[2] = static_cast<unsigned long long>(front_edges);
        }
        front_edges = s_front_edges;
        if (s_stop) break;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *a.levels_out = level;
}

__global__ void k_ms_seed(const int32_t* __restrict__ srcs, int32_t k, int32_t n, Bits* seen, Bits* front,
                          int32_t* out) {
    const int i = threadIdx.x;
    if (i < k) {
        const int32_t v = srcs[i];
        atomicOr(reinterpret_cast<unsigned long long*>(&seen[v].w[i >> 6]), 1ULL << (i & 63));
        atomicOr(reinterpret_cast<unsigned long long*>(&front[v].w[i >> 6]), 1ULL << (i & 63));
        out[static_cast<int64_t>(i) * n + v] = 0;
    }
}

// Depths from the level log: thread (v, word k) reads log[0..levels)[v].w[k]
// and, for each of the 64 sources of word k first reached within the log,
// writes out[source][v] (consecutive threads = consecutive v: coalesced).
__global__ void __launch_bounds__(256) k_ms_materialize(const Bits* __restrict__ log, int32_t levels, int32_t n,
                                                        int32_t kp, int32_t* __restrict__ out) {
    const int k = blockIdx.y;
    const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= n) return;
    uint64_t done = 0;
    for (int32_t L = 0; L < levels; ++L) {
        uint64_t x = log[static_cast<int64_t>(L) * n + v].w[k] & ~done;
        done |= x;
        while (x) {
            const int b = __ffsll(static_cast<unsigned long long>(x)) - 1;
            x &= x - 1;
            const int s = k * 64 + b;
            if (s < kp) out[static_cast<int64_t>(s) * n + v] = L + 1;
        }
    }
}

// Long in-lists of the in-CSR and their chunks (once per graph).
int build_ms_layout(ghet_graph* g, hipStream_t s) {
    MSLayout& L = g->ms;
    if (L.built) return GHET_OK;
    const std::vector<int32_t>& rp = g->h_in_rowptr;
    std::vector<int32_t> heavy;
    std::vector<int2> chunks;
    for (int32_t v = 0; v < g->n; ++v) {
        if (rp[v + 1] - rp[v] <= kPullHeavy) continue;
        heavy.push_back(v);
        for (int32_t e = rp[v]; e < rp[v + 1]; e += kChunkE) chunks.push_back(make_int2(v, e));
    }
    const size_t hb = (4 * heavy.size() + 255) & ~size_t(255);
    GHET_HIP(hipMalloc(&L.mem, hb + sizeof(int2) * chunks.size() + 256));
    L.heavy = static_cast<int32_t*>(L.mem);
    L.chunks = reinterpret_cast<int2*>(static_cast<char*>(L.mem) + hb);
    if (!heavy.empty()) GHET_HIP(hipMemcpyAsync(L.heavy, heavy.data(), 4 * heavy.size(), hipMemcpyHostToDevice, s));
    if (!chunks.empty())
        GHET_HIP(hipMemcpyAsync(L.chunks, chunks.data(), sizeof(int2) * chunks.size(), hipMemcpyHostToDevice, s));
    GHET_HIP(hipStreamSynchronize(s));
    L.num_heavy = static_cast<int32_t>(heavy.size());
    L.num_chunks = static_cast<int32_t>(chunks.size());
    L.built = true;
    return GHET_OK;
}

}  // namespace

int run_bfs_multi(ghet_graph* g, int32_t k, const int32_t* srcs, int32_t* depth_out, const ghet_opts* o,
                  ghet_run_stats* st) {
    const int32_t n = g->n;
    if (k < 0 || (k > 0 && !srcs)) return set_error(GHET_ERR_INVALID, "bad source list");
    for (int32_t i = 0; i < k; ++i)
        if (srcs[i] < 0 || srcs[i] >= n) return set_error(GHET_ERR_INVALID, "source %d out of range [0,%d)", srcs[i], n);
    if (!g->out.present() || !g->in.present()) return set_error(GHET_ERR_INVALID, "batched BFS needs out- and in-CSR");
    if (k == 0) return GHET_OK;
    hipStream_t s = pick_stream(g, o);
    GHET_CHECK(build_ms_layout(g, s));
    void* p;
    Bits *seen, *fa, *fb, *acc, *log;
    int32_t *out, *flags, *d_src, *d_levels;
    CtrSlot* slots;
    GHET_CHECK(g->ws.get(WS_MULTI_A, sizeof(Bits) * 4 * size_t(n), &p));
    seen = static_cast<Bits*>(p);
    fa = seen + n;
    fb = fa + n;
    acc = fb + n;
    GHET_CHECK(g->ws.get(WS_MULTI_B, sizeof(Bits) * kLogLevels * size_t(n), &p)); log = static_cast<Bits*>(p);
    const int32_t kmax = std::min(k, kPass);
    GHET_CHECK(g->ws.get(WS_MULTI_D, 4 * size_t(n) * kmax, &p)); out = static_cast<int32_t*>(p);
    // counters: [3][kSlots] per-level slots, [kSlots] totals, then flags[4], levels
    constexpr size_t kCtrBytes = sizeof(CtrSlot) * kSlots * 4 + 64;
    GHET_CHECK(g->ws.get(WS_COUNTERS, kCtrBytes, &p));
    slots = static_cast<CtrSlot*>(p);
    flags = reinterpret_cast<int32_t*>(slots + 4 * kSlots);
    d_levels = flags + 4;
    constexpr int kRecCap = 256;  // level records (slot 0 = start stamp)
    GHET_CHECK(g->ws.get(WS_MULTI_C, 4 * size_t(kPass) + 24 * kRecCap, &p));
    d_src = static_cast<int32_t*>(p);
    unsigned long long* d_rec = reinterpret_cast<unsigned long long*>(d_src + kPass);
    ghet_level_trace* trace = (o && o->size >= sizeof(ghet_opts)) ? o->trace : nullptr;
    const int trace_cap = trace ? o->trace_cap : 0;
    int trace_len = 0;
    int clock_khz = 0;
    if (trace) GHET_HIP(hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, g->device));
    int dev_blocks = 0, cus = 0;
    GHET_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&dev_blocks, reinterpret_cast<const void*>(k_msbfs), kBT, 0));
    GHET_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device));
    const unsigned grid = static_cast<unsigned>(std::max(1, std::min(dev_blocks, 8) * cus));
    const int64_t bits_words = sizeof(Bits) / 8 * 4 * int64_t(n);
    const bool have_opts = o && o->size >= sizeof(ghet_opts);
    const int64_t push_div = (have_opts && o->reserved[2] > 0) ? o->reserved[2] : 16;

    int64_t traversed = 0, scanned = 0, pushed = 0;
    int32_t max_levels_seen = 0;
    GHET_HIP(hipEventRecord(g->ev_begin, s));
    for (int32_t p0 = 0; p0 < k; p0 += kPass) {
        const int32_t kp = std::min(kPass, k - p0);
        GHET_CHECK(fill<uint64_t>(reinterpret_cast<uint64_t*>(seen), 0, bits_words, s));
        GHET_CHECK(fill<int32_t>(out, kIntInf, int64_t(n) * kp, s));
        GHET_CHECK(fill<uint64_t>(reinterpret_cast<uint64_t*>(slots), 0, kCtrBytes / 8, s));
        GHET_HIP(hipMemcpyAsync(d_src, srcs + p0, 4 * size_t(kp), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_ms_seed, dim3(1), dim3(kPass), 0, s, d_src, kp, n, seen, fa, out);
        MsArgs a;
        a.n = n;
        a.out_rowptr = g->out.rowptr;
        a.out_col = g->out.col;
        a.in_rowptr = g->in.rowptr;
        a.in_col = g->in.col;
        a.seen = seen;
        a.fa = fa;
        a.fb = fb;
        a.acc = acc;
        a.log = log;
        a.out = out;
        for (int w = 0; w < kWords; ++w) {
            const int lo = w * 64, hi = std::min(kp, lo + 64);
            a.full.w[w] = hi <= lo ? 0ULL : (hi - lo == 64 ? ~0ULL : ((1ULL << (hi - lo)) - 1));
        }
        a.flags = flags;
        a.slots = slots;
        // the seeds' distinct vertices form level 0's frontier
        std::vector<int32_t> uniq(srcs + p0, srcs + p0 + kp);
        std::sort(uniq.begin(), uniq.end());
        uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
        int64_t fe = 0;
        for (int32_t v : uniq) fe += g->h_out_rowptr[v + 1] - g->h_out_rowptr[v];
        for (int32_t i = 0; i < kp; ++i) traversed += g->h_out_rowptr[srcs[p0 + i] + 1] - g->h_out_rowptr[srcs[p0 + i]];
        a.front_edges0 = fe;
        a.push_limit = std::max<int64_t>(g->out.m / push_div, 1);
        a.max_levels = n + 1;
        a.levels_out = d_levels;
        a.lvl_rec = (trace && p0 == 0) ? d_rec : nullptr;
        a.lvl_cap = kRecCap;
        a.heavy = g->ms.heavy;
        a.num_heavy = g->ms.num_heavy;
        a.chunks = g->ms.chunks;
        a.num_chunks = g->ms.num_chunks;
        void* args[] = {&a};
        GHET_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_msbfs), dim3(grid), dim3(kBT), args, 0, s));
        GHET_HIP(hipMemcpyAsync(g->pinned, slots, kCtrBytes, hipMemcpyDeviceToHost, s));
        unsigned long long* h_rec = reinterpret_cast<unsigned long long*>(reinterpret_cast<char*>(g->pinned) + kCtrBytes);
        static_assert(kCtrBytes + 24 * kRecCap <= kPinnedBytes, "pinned buffer");
        if (a.lvl_rec) GHET_HIP(hipMemcpyAsync(h_rec, d_rec, 24 * kRecCap, hipMemcpyDeviceToHost, s));
        GHET_HIP(hipStreamSynchronize(s));
        const CtrSlot* hs = reinterpret_cast<const CtrSlot*>(g->pinned);
        const int32_t lv = reinterpret_cast<const int32_t*>(hs + 4 * kSlots)[4];
        max_levels_seen = std::max(max_levels_seen, lv);
        if (a.lvl_rec && clock_khz > 0) {
            // first pass: one record per level (mode, frontier out-edges / m, device ms)
            for (int32_t L = 0; L < lv && L + 1 < kRecCap && trace_len < trace_cap; ++L) {
                ghet_level_trace& t = trace[trace_len++];
                std::memset(&t, 0, sizeof(t));
                const unsigned long long* r = h_rec + 3 * (L + 1);
                t.mode = static_cast<int32_t>(r[1]);
                t.features[0] = static_cast<float>(static_cast<double>(r[2]) / std::max<int64_t>(g->out.m, 1));
                t.ms = static_cast<float>(static_cast<double>(r[0] - h_rec[3 * L]) / clock_khz);
            }
        }
        traversed += static_cast<int64_t>(slot_sum(hs + 3 * kSlots, 1));
        scanned += static_cast<int64_t>(slot_sum(hs + 3 * kSlots, 2));
        pushed += static_cast<int64_t>(slot_sum(hs + 3 * kSlots, 3));
        // the last level found nothing new; levels [0, lv) hold log entries
        const int32_t logged = std::min(lv, kLogLevels);
        if (logged > 0) {
            dim3 mg(static_cast<unsigned>(ceil_div(n, 256)), static_cast<unsigned>(ceil_div(kp, 64)));
            hipLaunchKernelGGL(k_ms_materialize, mg, dim3(256), 0, s, log, logged, n, kp, out);
        }
        if (depth_out)
            GHET_HIP(hipMemcpyAsync(depth_out + static_cast<int64_t>(p0) * n, out, 4 * size_t(n) * kp,
                                    hipMemcpyDeviceToHost, s));
    }
    GHET_HIP(hipEventRecord(g->ev_end, s));
    GHET_HIP(hipStreamSynchronize(s));
    if (st) {
        float ms = 0;
        GHET_HIP(hipEventElapsedTime(&ms, g->ev_begin, g->ev_end));
        st->device_ms = ms;
        st->levels = max_levels_seen;
        st->edges_traversed = traversed;
        st->extra[0] = grid;
        st->extra[1] = static_cast<double>(scanned);
        st->extra[2] = static_cast<double>(pushed);
        st->trace_len = trace_len;
    }
    return GHET_OK;
}

}  // namespace ghet
