// Triangle counting on gfx950: per vertex u, its list N(u) goes into a hash
// table in LDS, and the lists N(v) of its neighbours stream past it.
//
// Reference: Cuda_TC_PerNode (GPU_TC_csr.cuh:46-69) runs one block per vertex
// u and one thread per v in N(u), merging N(u) and N(v) and adding 1 to u, v
// and w for every common w (three int32 atomics); CPU Tri_Counting
// (CPU_TC.hpp:31-60) is the same loop serially.  On a symmetric loop-free
// graph each vertex receives 6 * (triangles containing it).
//
// Two forms with identical output:
//   general  the reference loop itself over the out-CSR, for any input;
//   DAG      when the out-CSR is symmetric and loop-free: orient every edge
//            from lower to higher (degree, id), count each triangle once as
//            (u, v, w) with u->v, u->w, v->w, add 1 to each, and report 6x.
// Counts are int64 (the reference's int32 can wrap at scale-24 hubs).
//
// Counting kernels.  Every (u, v, w) counted is found from v: v's list N(v)
// goes into a hash table in LDS, and for every u whose list holds v (the
// transpose's in-list of v) N(u) streams past the table.  A hit on w adds 1
// to w's slot counter, to u's count and to v's.  Streaming N(u) once per
// edge u -> v costs sum over u of |N(u)|^2 entries, against sum over v of
// |in(v)| |N(v)| for streaming N(v) past u's table: on the degree-ordered
// DAG of R-MAT the first is about 1.8 times smaller (2.5e9 against 4.4e9
// entries at scale 20).  In the DAG form each row is kept in the DAG's own
// order and each in-entry knows where v sits in N(u), so only the part of
// N(u) after v streams: N(v) holds only vertices after v, so nothing before
// it can hit, and the entries fall to sum over u of |N(u)| (|N(u)| - 1) / 2.
// When v's task is done, the slot counters go to the
// global counts with one 64-bit atomic per neighbour that was hit, v's total
// with one, and each u's with one: device-scope atomics per vertex and edge
// instead of per triangle corner (they go to the fabric: the XCDs' L2s are
// not coherent).
//   tasks       v and a run of its in-list (kInWave or kInBlock entries), so
//               a hub with a long in-list is shared by many waves;
//   k_tc_wave   tasks of v with up to kWaveMax entries: one wave each, a
//               table of at most 2 kWaveMax slots per wave, tasks taken
//               `vpt` at a time from an atomic counter;
//   k_tc_block  tasks of longer v: one 1024-thread block each, a table of up
//               to kBlockSlots slots filled kBlockSlots / 2 entries of N(v)
//               at a time.
// Tables are 4-slot buckets (multiplicative hash, linear probing over
// buckets), at most half full, so every probe ends at the key or at a bucket
// with a free slot.  The lists are those of a canonical CSR (no duplicate
// entries in a row).
#include "graph_internal.hpp"

#include <rocprim/rocprim.hpp>

namespace ghet {
namespace {

constexpr int kBT = 256;
constexpr int kWaveMax = 256;                // rows up to this length: k_tc_wave
constexpr int kWaveSlots = 2 * kWaveMax;     // table slots per wave
constexpr int kBlockBT = 1024;               // k_tc_block threads
constexpr int kBlockSlots = 8192;            // k_tc_block table slots (keys + counters: 64 KB, two blocks per CU)
constexpr int kInWave = 256;                 // in-list entries per wave task
constexpr int kInBlock = 4096;               // in-list entries per block task
constexpr int kStreamQ = 8;                  // list entries in flight per lane (4: 3 % slower at s24)

__device__ __forceinline__ int upper_bound_i32(const int32_t* a, int64_t len, int32_t key) {
    int64_t lo = 0, hi = len;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (a[mid] <= key) lo = mid + 1;
        else hi = mid;
    }
    return static_cast<int>(lo);
}

__device__ __forceinline__ bool contains(const int32_t* a, int32_t len, int32_t key) {
    int32_t lo = 0, hi = len;
    while (lo < hi) {
        int32_t mid = (lo + hi) >> 1;
        int32_t x = a[mid];
        if (x == key) return true;
        if (x < key) lo = mid + 1;
        else hi = mid;
    }
    return false;
}

// Flags any self-loop or any edge (u,v) without its reverse (v,u).
__global__ void k_tc_check(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col, int32_t n, int64_t m,
                           int32_t* bad) {
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < m;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int32_t u = upper_bound_i32(rowptr, static_cast<int64_t>(n) + 1, static_cast<int32_t>(e)) - 1;
        const int32_t v = col[e];
        if (u == v || !contains(col + rowptr[v], rowptr[v + 1] - rowptr[v], u)) atomicOr(bad, 1);
    }
}

__device__ __forceinline__ bool ranks_above(int32_t du, int32_t u, int32_t dv, int32_t v) {
    return dv > du || (dv == du && v > u);
}

// Pass 0 counts, pass 1 writes (order-preserving) the out-neighbours of u
// that rank above u.  One wave per vertex.
__global__ void __launch_bounds__(kBT) k_dag(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                             int32_t n, int32_t* __restrict__ dag_deg,
                                             const int32_t* __restrict__ dag_rowptr, int32_t* __restrict__ dag_col,
                                             int pass) {
    const int lane = lane_id();
    const int64_t wave_global = (static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x) >> 6;
    const int64_t wave_stride = (static_cast<int64_t>(gridDim.x) * kBT) >> 6;
    for (int64_t u = wave_global; u < n; u += wave_stride) {
        const int32_t b = rowptr[u], e = rowptr[u + 1], du = e - b;
        int32_t out = pass ? dag_rowptr[u] : 0;
        for (int32_t i0 = b; i0 < e; i0 += 64) {
            const int32_t i = i0 + lane;
            bool keep = false;
            int32_t v = 0;
            if (i < e) {
                v = col[i];
                keep = ranks_above(du, static_cast<int32_t>(u), rowptr[v + 1] - rowptr[v], v);
            }
            const uint64_t mask = __ballot(keep);
            if (pass && keep) dag_col[out + lanes_below(mask)] = v;
            out += __popcll(mask);
        }
        if (!pass && lane == 0) dag_deg[u] = out;
    }
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t tc_slot0(int32_t key, int shift) {
    return (static_cast<uint32_t>(key) * 2654435761u) >> shift;
}

// Tables are arrays of 4-slot buckets (one 128-bit LDS read each).  An
// insert takes the first free slot of its home bucket, else of the next
// bucket, so a bucket's slots fill in order: a lookup that meets a bucket
// with a free last slot has seen every key that could be there.
__device__ __forceinline__ void tc_insert(int32_t* keys, uint32_t bmask, int shift, int32_t key) {
    uint32_t b = tc_slot0(key, shift);
    for (;;) {
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (atomicCAS(&keys[4 * b + j], -1, key) == -1) return;
        b = (b + 1) & bmask;
    }
}

__device__ __forceinline__ int tc_find(const int32_t* keys, uint32_t bmask, int shift, int32_t key) {
    uint32_t b = tc_slot0(key, shift);
    for (;;) {
        const int4 q = reinterpret_cast<const int4*>(keys)[b];
        const int r = q.x == key ? 0 : q.y == key ? 1 : q.z == key ? 2 : q.w == key ? 3 : -1;
        if (r >= 0) return static_cast<int>(4 * b) + r;
        if (q.w < 0) return -1;
        b = (b + 1) & bmask;
    }
}

// Per-wave LDS of tc_stream: list starts and offsets of the batch, and
// each list's hits.
struct TcBatch {
    int32_t pre[65];   // exclusive prefix of the list lengths; pre[64] = total
    int32_t base[64];  // first entry of each list in col
    uint32_t hits[64];
};

// One list per lane (len 0: none), concatenated: the wave walks entries
// k = 0 .. total-1 of the concatenation kStreamQ per lane at a time, so
// neighbouring lanes load neighbouring entries of a list (coalesced) and
// kStreamQ loads are in flight per lane.  Each entry found in the table adds 1
// to its slot's counter and to its list's hits; returns the lane's own
// list's hits.
__device__ __forceinline__ uint32_t tc_stream(int32_t lb, int32_t len, const int32_t* __restrict__ col,
                                              const int32_t* keys, uint32_t* cnts, uint32_t mask, int shift,
                                              TcBatch& bt) {
    const int lane = lane_id();
    int32_t incl = len;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int32_t y = __shfl_up(incl, off);
        if (lane >= off) incl += y;
    }
    const int32_t total = __shfl(incl, 63);
    if (total == 0) return 0;
    bt.pre[lane] = incl - len;
    bt.base[lane] = lb;
    bt.hits[lane] = 0;
    if (lane == 0) bt.pre[64] = total;
    wave_lds_sync();
    int owner = 0;  // list of entry k: the last i with pre[i] <= k (k only grows)
    for (int32_t k0 = 0; k0 < total; k0 += kStreamQ * 64) {
        int32_t w[kStreamQ];
        int own[kStreamQ];
#pragma unroll
        for (int j = 0; j < kStreamQ; ++j) {
            const int32_t k = k0 + j * 64 + lane;
            w[j] = -1;
            own[j] = 0;
            if (k < total) {
                while (bt.pre[owner + 1] <= k) ++owner;
                own[j] = owner;
                w[j] = col[bt.base[owner] + (k - bt.pre[owner])];
            }
        }
#pragma unroll
        for (int j = 0; j < kStreamQ; ++j) {
            if (w[j] >= 0) {
                const int sl = tc_find(keys, mask, shift, w[j]);
                if (sl >= 0) {
                    atomicAdd(&cnts[sl], 1u);
                    atomicAdd(&bt.hits[own[j]], 1u);
                }
            }
        }
    }
    wave_lds_sync();
    const uint32_t c = bt.hits[lane];
    wave_lds_sync();  // the next batch rewrites the arrays
    return c;
}

// A task is a vertex v and a run of at most kInWave (wave) or kInBlock
// (block) entries of its in-list, from `first`; tasks are taken `vpt` at a
// time from *task_ctr (zero at launch).
struct alignas(8) TcTask {
    int32_t v, first;
};

// The part of N(u) an in-entry j of v streams: all of it, or (in_pos given,
// DAG form) the entries after v, which sits at in_pos[j] in u's row.
__device__ __forceinline__ void in_entry_list(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ in_col,
                                              const int32_t* __restrict__ in_pos, int32_t j, int32_t* u, int32_t* ub,
                                              int32_t* len) {
    *u = in_col[j];
    const int32_t end = rowptr[*u + 1];
    *ub = in_pos ? in_pos[j] + 1 : rowptr[*u];
    *len = end - *ub;
}

// Tasks of vertices with |N(v)| <= kWaveMax: one wave each.  The wave
// builds v's table, then every u of the task's in-list run streams N(u)
// (or its part after v) past it: c hits add c to u (one atomic per u with
// hits) and to v.
__global__ void __launch_bounds__(kBT) k_tc_wave(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                                 const int32_t* __restrict__ in_rowptr,
                                                 const int32_t* __restrict__ in_col, const int32_t* __restrict__ in_pos,
                                                 const TcTask* __restrict__ tasks,
                                                 int32_t num_tasks, int vpt, int* task_ctr, unsigned long long* cnt) {
    __shared__ __align__(16) int32_t s_keys[kBT / 64][kWaveSlots];
    __shared__ uint32_t s_cnts[kBT / 64][kWaveSlots];
    __shared__ TcBatch s_batch[kBT / 64];
    const int lane = lane_id(), wave = threadIdx.x >> 6;
    int32_t* keys = s_keys[wave];
    uint32_t* cnts = s_cnts[wave];
    for (;;) {
        int32_t t0 = 0;
        if (lane == 0) t0 = atomicAdd(task_ctr, vpt);
        t0 = __shfl(t0, 0);
        if (t0 >= num_tasks) break;
        const int32_t t1 = num_tasks - t0 < vpt ? num_tasks : t0 + vpt;
        for (int32_t t = t0; t < t1; ++t) {
            const TcTask task = tasks[t];
            const int32_t v = task.v, b = rowptr[v], d = rowptr[v + 1] - b;
            const int32_t j0 = in_rowptr[v] + task.first;
            const int32_t j1 = min(in_rowptr[v + 1], j0 + kInWave);
            int tbits = 6;  // 4 d slots (2 d at most for the longest rows)
            while ((1 << tbits) < 4 * d && (1 << tbits) < kWaveSlots) ++tbits;
            const uint32_t T = 1u << tbits, mask = (T >> 2) - 1;
            const int shift = 32 - (tbits - 2);
            for (uint32_t sl = lane; sl < T; sl += 64) {
                keys[sl] = -1;
                cnts[sl] = 0;
            }
            wave_lds_sync();
            for (int32_t i = lane; i < d; i += 64) tc_insert(keys, mask, shift, col[b + i]);
            wave_lds_sync();
            uint32_t tot = 0;
            for (int32_t jb = j0; jb < j1; jb += 64) {
                const int32_t j = jb + lane;
                int32_t u = -1, ub = 0, du = 0;
                if (j < j1) in_entry_list(rowptr, in_col, in_pos, j, &u, &ub, &du);
                const uint32_t c = tc_stream(ub, du, col, keys, cnts, mask, shift, s_batch[wave]);
                if (c) {
                    atomicAdd(&cnt[u], static_cast<unsigned long long>(c));
                    tot += c;
                }
            }
            wave_lds_sync();
            const unsigned long long total = wave_sum(static_cast<unsigned long long>(tot));
            if (lane == 0 && total) atomicAdd(&cnt[v], total);
            for (uint32_t sl = lane; sl < T; sl += 64) {
                const uint32_t k = cnts[sl];
                if (k) atomicAdd(&cnt[keys[sl]], static_cast<unsigned long long>(k));
            }
            wave_lds_sync();  // the next task clears the table
        }
    }
}

// Tasks of vertices with |N(v)| > kWaveMax: one block each; v's table is
// filled kBlockSlots / 2 entries at a time, and the task's in-list run
// streams past each fill.
__global__ void __launch_bounds__(kBlockBT, 8) k_tc_block(const int32_t* __restrict__ rowptr,
                                                          const int32_t* __restrict__ col,
                                                          const int32_t* __restrict__ in_rowptr,
                                                          const int32_t* __restrict__ in_col,
                                                          const int32_t* __restrict__ in_pos,
                                                          const TcTask* __restrict__ tasks, int32_t num_tasks,
                                                          unsigned long long* cnt) {
    __shared__ __align__(16) int32_t keys[kBlockSlots];
    __shared__ uint32_t cnts[kBlockSlots];
    __shared__ TcBatch s_batch[kBlockBT / 64];
    __shared__ unsigned long long s_red[kBlockBT / 64];
    const int tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
    for (int32_t t = blockIdx.x; t < num_tasks; t += gridDim.x) {
        const TcTask task = tasks[t];
        const int32_t v = task.v, b = rowptr[v], d = rowptr[v + 1] - b;
        const int32_t j0 = in_rowptr[v] + task.first;
        const int32_t j1 = min(in_rowptr[v + 1], j0 + kInBlock);
        int tbits = 6;  // 4 d slots, at most kBlockSlots, filled to half at most
        while ((1 << tbits) < 4 * d && (1 << tbits) < kBlockSlots) ++tbits;
        const uint32_t T = 1u << tbits, mask = (T >> 2) - 1;
        const int shift = 32 - (tbits - 2);
        const int32_t cap = static_cast<int32_t>(T / 2);
        unsigned long long tot = 0;
        for (int32_t c0 = 0; c0 < d; c0 += cap) {
            const int32_t c1 = d - c0 < cap ? d : c0 + cap;
            for (uint32_t sl = tid; sl < T; sl += kBlockBT) {
                keys[sl] = -1;
                cnts[sl] = 0;
            }
            __syncthreads();
            for (int32_t i = c0 + tid; i < c1; i += kBlockBT) tc_insert(keys, mask, shift, col[b + i]);
            __syncthreads();
            for (int32_t jb = j0 + wave * 64; jb < j1; jb += kBlockBT) {
                const int32_t j = jb + lane;
                int32_t u = -1, ub = 0, du = 0;
                if (j < j1) in_entry_list(rowptr, in_col, in_pos, j, &u, &ub, &du);
                const uint32_t c = tc_stream(ub, du, col, keys, cnts, mask, shift, s_batch[wave]);
                if (c) {
                    atomicAdd(&cnt[u], static_cast<unsigned long long>(c));
                    tot += c;
                }
            }
            __syncthreads();
            for (uint32_t sl = tid; sl < T; sl += kBlockBT) {
                const uint32_t k = cnts[sl];
                if (k) atomicAdd(&cnt[keys[sl]], static_cast<unsigned long long>(k));
            }
            __syncthreads();
        }
        tot = block_sum_u64<kBlockBT>(tot, s_red);
        if (tid == 0 && tot) atomicAdd(&cnt[v], tot);
    }
}

// Transpose of the counted CSR: the source of every edge, in row order.
// Sort keys putting a DAG row in the DAG's order: (degree, id) of each
// entry, the degree from the symmetric CSR the DAG was oriented by.
__global__ void k_dag_order_keys(const int32_t* __restrict__ col, int64_t m, const int32_t* __restrict__ rowptr,
                                 unsigned long long* __restrict__ key) {
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < m;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int32_t w = col[e];
        key[e] = (static_cast<unsigned long long>(rowptr[w + 1] - rowptr[w]) << 32) | static_cast<uint32_t>(w);
    }
}
__global__ void k_dag_order_ids(const unsigned long long* __restrict__ key, int64_t m, int32_t* __restrict__ col) {
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < m;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x)
        col[e] = static_cast<int32_t>(key[e] & 0xFFFFFFFFull);
}

// Edge indices 0 .. m-1 (the values of the transpose's sort).
__global__ void k_tc_iota(int32_t* __restrict__ v, int64_t m) {
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < m;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x)
        v[e] = static_cast<int32_t>(e);
}

// In-entries from the sorted edge indices: the source of each edge and its
// position in the source's row.
__global__ void k_tc_in_fill(const int32_t* __restrict__ sorted_e, const int32_t* __restrict__ src, int64_t m,
                             int32_t* __restrict__ in_col, int32_t* __restrict__ in_pos) {
    for (int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; j < m;
         j += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int32_t e = sorted_e[j];
        in_col[j] = src[e];
        in_pos[j] = e;
    }
}

__global__ void __launch_bounds__(kBT) k_tc_sources(const int32_t* __restrict__ rowptr, int32_t n,
                                                    int32_t* __restrict__ src) {
    const int lane = lane_id();
    const int64_t wave_global = (static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x) >> 6;
    const int64_t wave_stride = (static_cast<int64_t>(gridDim.x) * kBT) >> 6;
    for (int64_t u = wave_global; u < n; u += wave_stride)
        for (int32_t e = rowptr[u] + lane; e < rowptr[u + 1]; e += 64) src[e] = static_cast<int32_t>(u);
}

// off[k] = first position of key k in the sorted keys; off[num_keys] = count.
__global__ void k_tc_key_offsets(const int32_t* __restrict__ key, int64_t count, int32_t num_keys,
                                 int32_t* __restrict__ off) {
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e <= count;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t prev = e > 0 ? static_cast<int64_t>(key[e - 1]) : -1;
        const int64_t cur = e < count ? static_cast<int64_t>(key[e]) : static_cast<int64_t>(num_keys);
        for (int64_t k = prev + 1; k <= cur; ++k) off[k] = static_cast<int32_t>(e);
    }
}

// Tasks per vertex: none unless v has out- and in-entries (a triangle needs
// w in N(v) and u with v in N(u)); then its in-list in runs of kInWave
// (table fits a wave) or kInBlock.
__global__ void k_tc_task_count(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ in_rowptr, int32_t n,
                                int32_t* __restrict__ nw, int32_t* __restrict__ nb) {
    for (int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < n;
         v += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int32_t d = rowptr[v + 1] - rowptr[v], di = in_rowptr[v + 1] - in_rowptr[v];
        const bool any = d > 0 && di > 0;
        nw[v] = any && d <= kWaveMax ? (di + kInWave - 1) / kInWave : 0;
        nb[v] = any && d > kWaveMax ? (di + kInBlock - 1) / kInBlock : 0;
    }
}
__global__ void k_tc_task_fill(const int32_t* __restrict__ nw, const int32_t* __restrict__ ow,
                               const int32_t* __restrict__ nb, const int32_t* __restrict__ ob, int32_t n,
                               TcTask* __restrict__ wt, TcTask* __restrict__ bt) {
    for (int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < n;
         v += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        for (int32_t k = 0; k < nw[v]; ++k) wt[ow[v] + k] = TcTask{static_cast<int32_t>(v), k * kInWave};
        for (int32_t k = 0; k < nb[v]; ++k) bt[ob[v] + k] = TcTask{static_cast<int32_t>(v), k * kInBlock};
    }
}

// Intersection bytes of the oriented CSR: over its edges (u, v), 4 bytes per
// entry of both lists, 4 (d+(u) + d+(v)); per-block sums into spread slots.
__global__ void __launch_bounds__(kBT) k_isect_bytes(const int32_t* __restrict__ rowptr,
                                                     const int32_t* __restrict__ col, int32_t n, CtrSlot* slots) {
    __shared__ unsigned long long s_red[kBT / 64];
    unsigned long long x = 0;
    for (int64_t u = static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x; u < n;
         u += static_cast<int64_t>(gridDim.x) * kBT) {
        const int32_t b = rowptr[u], e = rowptr[u + 1];
        const unsigned long long du = static_cast<unsigned long long>(e - b);
        for (int32_t i = b; i < e; ++i) x += du + static_cast<unsigned long long>(rowptr[col[i] + 1] - rowptr[col[i]]);
    }
    x = block_sum_u64<kBT>(x, s_red);
    if (threadIdx.x == 0) slot_add(slots, 0, 4 * x);
}

__global__ void k_scale_i64(unsigned long long* c, int32_t n, unsigned long long f) {
    for (int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < n;
         v += static_cast<int64_t>(gridDim.x) * blockDim.x)
        c[v] *= f;
}

unsigned grid_cap(int64_t blocks) {
    return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(blocks, 65535)));
}

int build_tc_layout(ghet_graph* g, hipStream_t s) {
    TCLayout& L = g->tc;
    if (L.built) return GHET_OK;
    const int32_t n = g->n;
    const int64_t m = g->out.m;
    void* p;
    GHET_CHECK(g->ws.get(WS_COUNTERS, 256, &p));
    int32_t* bad = static_cast<int32_t*>(p);
    GHET_CHECK(fill<int32_t>(bad, 0, 1, s));
    if (m > 0)
        hipLaunchKernelGGL(k_tc_check, dim3(grid_cap(ceil_div(m, 256))), dim3(256), 0, s, g->out.rowptr, g->out.col, n,
                           m, bad);
    GHET_HIP(hipMemcpyAsync(g->pinned, bad, 4, hipMemcpyDeviceToHost, s));
    GHET_HIP(hipStreamSynchronize(s));
    L.symmetric = (g->pinned[0] == 0);
    if (L.symmetric && n > 0) {
        const int64_t m_dag = m / 2;
        size_t bytes = ((4 * (size_t(n) + 1) + 255) & ~size_t(255)) + ((4 * size_t(m_dag) + 255) & ~size_t(255)) + 256;
        GHET_HIP(hipMalloc(&L.mem, bytes));
        int32_t* drp = static_cast<int32_t*>(L.mem);
        int32_t* dcol = reinterpret_cast<int32_t*>(static_cast<char*>(L.mem) + ((4 * (size_t(n) + 1) + 255) & ~size_t(255)));
        int32_t *deg, *aux;
        GHET_CHECK(g->ws.get(WS_SCAN, 4 * (size_t(n) + 1), &p)); deg = static_cast<int32_t*>(p);
        GHET_CHECK(g->ws.get(WS_SCAN_AUX, 4 * size_t(scan_aux_ints(n)), &p)); aux = static_cast<int32_t*>(p);
        const unsigned gw = grid_cap(ceil_div(n, kBT / 64));
        hipLaunchKernelGGL(k_dag, dim3(gw), dim3(kBT), 0, s, g->out.rowptr, g->out.col, n, deg, drp, dcol, 0);
        GHET_CHECK(exclusive_scan_i32(deg, drp, n, aux, s));
        hipLaunchKernelGGL(k_dag, dim3(gw), dim3(kBT), 0, s, g->out.rowptr, g->out.col, n, deg, drp, dcol, 1);
        GHET_HIP(hipGetLastError());
        if (m_dag > 0) {
            // each row in the DAG's order, (degree, id) ascending: a
            // segmented radix sort of 64-bit keys, one segment per row
            DevTemp t;
            unsigned long long* k0 = t.get<unsigned long long>(size_t(m_dag));
            unsigned long long* k1 = t.get<unsigned long long>(size_t(m_dag));
            GHET_HIP(t.err);
            const unsigned ge = grid_cap(std::min<int64_t>(ceil_div(m_dag, 256), 65535));
            hipLaunchKernelGGL(k_dag_order_keys, dim3(ge), dim3(256), 0, s, dcol, m_dag, g->out.rowptr, k0);
            size_t sb = 0;
            GHET_HIP(rocprim::segmented_radix_sort_keys(nullptr, sb, k0, k1, static_cast<unsigned>(m_dag),
                                                        static_cast<unsigned>(n), drp, drp + 1, 0, 64, s));
            void* st = t.get<char>(sb);
            GHET_HIP(t.err);
            GHET_HIP(rocprim::segmented_radix_sort_keys(st, sb, k0, k1, static_cast<unsigned>(m_dag),
                                                        static_cast<unsigned>(n), drp, drp + 1, 0, 64, s));
            hipLaunchKernelGGL(k_dag_order_ids, dim3(ge), dim3(256), 0, s, k1, m_dag, dcol);
            GHET_HIP(hipGetLastError());
        }
        GHET_HIP(hipStreamSynchronize(s));
        L.dag = {n, m_dag, drp, dcol, nullptr};
    }
    {
        // transpose of the counted CSR (its in-lists, sources ascending) and
        // the task lists of both kernels
        const DevCSR& X = L.symmetric ? L.dag : g->out;
        const int64_t mx = X.m;
        if (mx >= INT32_MAX) return set_error(GHET_ERR_UNSUPPORTED, "TC: %lld edges", static_cast<long long>(mx));
        int32_t *ka = nullptr, *kb = nullptr, *va = nullptr, *vb = nullptr, *nw = nullptr, *nbk = nullptr, *ow = nullptr,
                *ob = nullptr, *aux = nullptr, *srcs = nullptr;
        void* tmp = nullptr;
        size_t tb = 0;
        hipError_t e = hipSuccess;
        auto get = [&](int32_t** q, size_t cnt) {
            if (e == hipSuccess) e = hipMalloc(q, std::max<size_t>(4 * cnt, 256));
        };
        get(&ka, size_t(mx));
        get(&kb, size_t(mx));
        get(&va, size_t(mx));
        get(&vb, size_t(mx));
        get(&srcs, size_t(mx));
        get(&nw, size_t(n));
        get(&nbk, size_t(n));
        get(&ow, size_t(n) + 1);
        get(&ob, size_t(n) + 1);
        get(&aux, size_t(scan_aux_ints(n)));
        const size_t col_b = (4 * size_t(mx) + 255) & ~size_t(255);
        const size_t in_b = ((4 * (size_t(n) + 1) + 255) & ~size_t(255)) + 2 * col_b;
        if (e == hipSuccess) e = hipMalloc(&L.in_mem, in_b + 256);
        if (e == hipSuccess) {
            L.in_rowptr = static_cast<int32_t*>(L.in_mem);
            L.in_col = reinterpret_cast<int32_t*>(static_cast<char*>(L.in_mem) + ((4 * (size_t(n) + 1) + 255) & ~size_t(255)));
            L.in_pos = reinterpret_cast<int32_t*>(reinterpret_cast<char*>(L.in_col) + col_b);
            L.task_ctr = reinterpret_cast<int*>(static_cast<char*>(L.in_mem) + in_b);
            if (mx > 0) {
                // sort the edge indices by target (stable: each in-list in
                // source order), then read each edge's source and position
                hipLaunchKernelGGL(k_tc_sources, dim3(grid_cap(ceil_div(n, kBT / 64))), dim3(kBT), 0, s, X.rowptr, n, srcs);
                hipLaunchKernelGGL(k_tc_iota, dim3(grid_cap(std::min<int64_t>(ceil_div(mx, 256), 65535))), dim3(256), 0, s,
                                   va, mx);
                e = hipMemcpyAsync(ka, X.col, 4 * size_t(mx), hipMemcpyDeviceToDevice, s);
            }
        }
        unsigned bits = 1;
        while ((int64_t(1) << bits) < n) ++bits;
        rocprim::double_buffer<int32_t> kd(ka, kb), vd(va, vb);
        if (e == hipSuccess && mx > 0) e = rocprim::radix_sort_pairs(nullptr, tb, kd, vd, static_cast<unsigned>(mx), 0, bits, s);
        if (e == hipSuccess && mx > 0) e = hipMalloc(&tmp, std::max<size_t>(tb, 256));
        if (e == hipSuccess && mx > 0) e = rocprim::radix_sort_pairs(tmp, tb, kd, vd, static_cast<unsigned>(mx), 0, bits, s);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_tc_key_offsets, dim3(grid_cap(std::min<int64_t>(ceil_div(mx + 1, 256), 65535))), dim3(256), 0,
                               s, kd.current(), mx, n, L.in_rowptr);
            if (mx > 0) {
                hipLaunchKernelGGL(k_tc_in_fill, dim3(grid_cap(std::min<int64_t>(ceil_div(mx, 256), 65535))), dim3(256), 0,
                                   s, vd.current(), srcs, mx, L.in_col, L.in_pos);
                e = hipGetLastError();
            }
        }
        if (e == hipSuccess) {
            const unsigned gn = grid_cap(std::min<int64_t>(ceil_div(n, 256), 4096));
            hipLaunchKernelGGL(k_tc_task_count, dim3(gn), dim3(256), 0, s, X.rowptr, L.in_rowptr, n, nw, nbk);
            e = hipGetLastError();
        }
        if (e == hipSuccess && exclusive_scan_i32(nw, ow, n, aux, s) != GHET_OK) e = hipErrorUnknown;
        if (e == hipSuccess && exclusive_scan_i32(nbk, ob, n, aux, s) != GHET_OK) e = hipErrorUnknown;
        int32_t totals[2] = {0, 0};
        if (e == hipSuccess) e = hipMemcpyAsync(totals, ow + n, 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipMemcpyAsync(totals + 1, ob + n, 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e == hipSuccess) {
            L.num_wave_tasks = totals[0];
            L.num_block_tasks = totals[1];
            e = hipMalloc(&L.task_mem, sizeof(TcTask) * (size_t(totals[0]) + size_t(totals[1])) + 256);
        }
        if (e == hipSuccess) {
            L.wave_tasks = L.task_mem;
            L.block_tasks = static_cast<TcTask*>(L.task_mem) + totals[0];
            const unsigned gn = grid_cap(std::min<int64_t>(ceil_div(n, 256), 4096));
            hipLaunchKernelGGL(k_tc_task_fill, dim3(gn), dim3(256), 0, s, nw, ow, nbk, ob, n,
                               static_cast<TcTask*>(L.wave_tasks), static_cast<TcTask*>(L.block_tasks));
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        for (void* q : {static_cast<void*>(ka), static_cast<void*>(kb), static_cast<void*>(va), static_cast<void*>(vb),
                        static_cast<void*>(nw), static_cast<void*>(nbk), static_cast<void*>(ow), static_cast<void*>(ob),
                        static_cast<void*>(aux), static_cast<void*>(srcs), tmp})
            if (q) hipFree(q);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return set_error(GHET_ERR_HIP, "TC layout: %s", hipGetErrorString(e));
        }
    }
    {
        // the reported intersection-bytes figure of the CSR the count reads
        const DevCSR& X = L.symmetric ? L.dag : g->out;
        GHET_CHECK(g->ws.get(WS_COUNTERS, sizeof(CtrSlot) * kSlots, &p));
        CtrSlot* slots = static_cast<CtrSlot*>(p);
        GHET_CHECK(fill<uint64_t>(reinterpret_cast<uint64_t*>(slots), 0, sizeof(CtrSlot) / 8 * kSlots, s));
        if (n > 0)
            hipLaunchKernelGGL(k_isect_bytes, dim3(grid_cap(std::min<int64_t>(ceil_div(n, kBT), 4096))), dim3(kBT), 0, s,
                               X.rowptr, X.col, n, slots);
        GHET_HIP(hipGetLastError());
        GHET_HIP(hipMemcpyAsync(g->pinned, slots, sizeof(CtrSlot) * kSlots, hipMemcpyDeviceToHost, s));
        GHET_HIP(hipStreamSynchronize(s));
        L.isect_bytes = static_cast<int64_t>(slot_sum(reinterpret_cast<const CtrSlot*>(g->pinned), 0));
    }
    L.built = true;
    return GHET_OK;
}

}  // namespace

int tc_prepare(ghet_graph* g, hipStream_t s) { return build_tc_layout(g, s); }

// Enqueues the counting (fill, count, 6x scale) with no host synchronisation,
// so it can be captured into a hipGraph.  vpt: tasks per grab of k_tc_wave;
// waves_per_cu: its grid size in resident waves per CU.
int enqueue_tc(ghet_graph* g, unsigned long long* cnt, int vpt, int waves_per_cu, hipStream_t s) {
    const int32_t n = g->n;
    const TCLayout& L = g->tc;
    const DevCSR& X = L.symmetric ? L.dag : g->out;
    const int32_t* in_pos = L.symmetric ? L.in_pos : nullptr;  // the DAG streams only past v
    GHET_CHECK(fill<uint64_t>(reinterpret_cast<uint64_t*>(cnt), 0ULL, n, s));
    if (X.m > 0) {
        int cus = 0;
        GHET_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device));
        if (L.num_block_tasks > 0)
            hipLaunchKernelGGL(k_tc_block, dim3(static_cast<unsigned>(std::min(L.num_block_tasks, 65535))), dim3(kBlockBT),
                               0, s, X.rowptr, X.col, L.in_rowptr, L.in_col, in_pos,
                               static_cast<const TcTask*>(L.block_tasks), L.num_block_tasks, cnt);
        if (L.num_wave_tasks > 0) {
            GHET_CHECK(fill<int32_t>(L.task_ctr, 0, 1, s));
            const int64_t grabs = ceil_div(L.num_wave_tasks, std::max(vpt, 1));
            const int64_t cap = int64_t(std::max(cus, 1)) * waves_per_cu / (kBT / 64);
            hipLaunchKernelGGL(k_tc_wave, dim3(grid_cap(std::min<int64_t>(ceil_div(grabs, kBT / 64), cap))), dim3(kBT), 0,
                               s, X.rowptr, X.col, L.in_rowptr, L.in_col, in_pos,
                               static_cast<const TcTask*>(L.wave_tasks), L.num_wave_tasks, std::max(vpt, 1), L.task_ctr,
                               cnt);
        }
    }
    if (L.symmetric)
        hipLaunchKernelGGL(k_scale_i64, dim3(grid_cap(std::min<int64_t>(ceil_div(n, 256), 8192))), dim3(256), 0, s, cnt,
                           n, 6ULL);
    GHET_HIP(hipGetLastError());
    return GHET_OK;
}

int run_tc(ghet_graph* g, int64_t* count_out, const ghet_opts* o, ghet_run_stats* st) {
    const int32_t n = g->n;
    if (!g->out.present()) return set_error(GHET_ERR_INVALID, "graph has no out-CSR");
    if (n == 0) return GHET_OK;
    hipStream_t s = pick_stream(g, o);
    GHET_CHECK(build_tc_layout(g, s));
    const TCLayout& L = g->tc;
    const DevCSR& X = L.symmetric ? L.dag : g->out;
    void* p;
    GHET_CHECK(g->ws.get(WS_TC_COUNT, 8 * size_t(n), &p));
    auto* cnt = static_cast<unsigned long long*>(p);
    GHET_HIP(hipEventRecord(g->ev_begin, s));
    GHET_CHECK(enqueue_tc(g, cnt, g->tune.tc_ept, g->tune.tc_waves_per_cu, s));
    GHET_HIP(hipEventRecord(g->ev_end, s));
    GHET_HIP(hipGetLastError());
    if (count_out) GHET_CHECK(copy_out(g, count_out, cnt, 8 * size_t(n), s));
    GHET_HIP(hipStreamSynchronize(s));
    if (st) {
        float ms = 0;
        GHET_HIP(hipEventElapsedTime(&ms, g->ev_begin, g->ev_end));
        st->device_ms = ms;
        st->hot_kernel_ms = ms;
        st->hot_kernel_launches = 1;
        st->levels = 1;
        st->extra[0] = L.symmetric ? 1.0 : 0.0;
        st->extra[1] = static_cast<double>(X.m);
        st->extra[2] = static_cast<double>(L.isect_bytes);  // 4 (d+(u) + d+(v)) over the counted edges
    }
    return GHET_OK;
}

}  // namespace ghet
