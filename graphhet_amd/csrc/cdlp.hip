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
//   d <= 32      one lane, labels staged in LDS, pairwise counting;
//   d <= 256     one wave, labels in LDS, each lane counts its labels;
//   d <= 2048    one workgroup, open-addressing hash table in LDS (4096 slots);
//   d > 2048     one workgroup per 4096-edge chunk counts in an LDS hash table,
//                then merges (label, count) pairs into the vertex's global
//                hash table (2^ceil(log2(2d)) slots); one workgroup takes the
//                argmax and clears the table for the next iteration.
// Winner = max count, ties to the smallest label: the maximum of the key
// (count << 32) | (0xFFFFFFFF - label).  The first iteration, from identity
// labels, has a closed form and skips the counting (k_cdlp_first_*).
#include "graph_internal.hpp"

#include <rocprim/rocprim.hpp>

namespace ghet {
namespace {

constexpr int kBT = 256;
constexpr int kSmall = 32, kWave = 256, kLds = 2048, kLdsSlots = 4096;
constexpr int kHeavyChunk = 4096;  // edges per heavy-vertex chunk
constexpr int kHeavySlots = 8192;  // LDS hash slots of a chunk (load <= 1/2)
constexpr int32_t kEmpty = -1;

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

__global__ void __launch_bounds__(kBT) k_cdlp_small(const int32_t* __restrict__ list, int32_t count,
                                                    const int32_t* __restrict__ rowptr,
                                                    const int32_t* __restrict__ col, const int32_t* __restrict__ lab,
                                                    int32_t* __restrict__ out) {
    __shared__ int32_t s[kBT * kSmall];
    int32_t* mine = s + threadIdx.x * kSmall;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x; i < count;
         i += static_cast<int64_t>(gridDim.x) * kBT) {
        const int32_t v = list[i];
        const int32_t b = rowptr[v], d = rowptr[v + 1] - b;
        for (int j = 0; j < d; ++j) mine[j] = lab[col[b + j]];
        uint64_t best = 0;
        for (int j = 0; j < d; ++j) {
            const int32_t x = mine[j];
            uint32_t c = 0;
            for (int k = 0; k < d; ++k) c += (mine[k] == x);
            const uint64_t key = pack(c, x);
            best = key > best ? key : best;
        }
        out[v] = unpack_label(best);
    }
}

__global__ void __launch_bounds__(kBT) k_cdlp_wave(const int32_t* __restrict__ list, int32_t count,
                                                   const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                                   const int32_t* __restrict__ lab, int32_t* __restrict__ out) {
    __shared__ int32_t s[(kBT / 64) * kWave];
    const int lane = lane_id();
    int32_t* mine = s + (threadIdx.x >> 6) * kWave;
    const int64_t wave_global = (static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x) >> 6;
    const int64_t wave_stride = (static_cast<int64_t>(gridDim.x) * kBT) >> 6;
    for (int64_t i = wave_global; i < count; i += wave_stride) {
        const int32_t v = list[i];
        const int32_t b = rowptr[v], d = rowptr[v + 1] - b;
        for (int j = lane; j < d; j += 64) mine[j] = lab[col[b + j]];
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        uint64_t best = 0;
        for (int j = lane; j < d; j += 64) {
            const int32_t x = mine[j];
            uint32_t c = 0;
            for (int k = 0; k < d; ++k) c += (mine[k] == x);
            const uint64_t key = pack(c, x);
            best = key > best ? key : best;
        }
        best = wave_max(best);
        if (lane == 0) out[v] = unpack_label(best);
        __builtin_amdgcn_wave_barrier();
    }
}

__global__ void __launch_bounds__(kBT) k_cdlp_lds(const int32_t* __restrict__ list, int32_t count,
                                                  const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                                  const int32_t* __restrict__ lab, int32_t* __restrict__ out) {
    __shared__ int32_t keys[kLdsSlots];
    __shared__ uint32_t cnts[kLdsSlots];
    __shared__ uint64_t s_red[kBT / 64];
    for (int64_t i = blockIdx.x; i < count; i += gridDim.x) {
        const int32_t v = list[i];
        const int32_t b = rowptr[v], d = rowptr[v + 1] - b;
        for (int j = threadIdx.x; j < kLdsSlots; j += kBT) {
            keys[j] = kEmpty;
            cnts[j] = 0;
        }
        __syncthreads();
        for (int j = threadIdx.x; j < d; j += kBT) {
            const int32_t x = lab[col[b + j]];
            uint32_t h = hash_slot(x, kLdsSlots - 1);
            for (;;) {
                int32_t prev = atomicCAS(&keys[h], kEmpty, x);
                if (prev == kEmpty || prev == x) {
                    atomicAdd(&cnts[h], 1u);
                    break;
                }
                h = (h + 1) & (kLdsSlots - 1);
            }
        }
        __syncthreads();
        uint64_t best = 0;
        for (int j = threadIdx.x; j < kLdsSlots; j += kBT) {
            if (keys[j] != kEmpty) {
                const uint64_t key = pack(cnts[j], keys[j]);
                best = key > best ? key : best;
            }
        }
        best = wave_max(best);
        if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = best;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t t = 0;
            for (int w = 0; w < kBT / 64; ++w) t = s_red[w] > t ? s_red[w] : t;
            out[v] = unpack_label(t);
        }
        __syncthreads();
    }
}

// Heavy vertices: chunk c counts the labels of edges [beg, beg+kHeavyChunk)
// of heavy vertex hv in an LDS hash table first, then adds one (label,
// count) pair per distinct label to the vertex's global table.  Once labels
// converge most of a hub's neighbours share a few labels; counting edge by
// edge in global memory would pile every one of them onto one counter.
__global__ void __launch_bounds__(kBT) k_cdlp_heavy_count(
    const int32_t* __restrict__ list, const int64_t* __restrict__ base, const int32_t* __restrict__ chunk_vtx,
    const int32_t* __restrict__ chunk_beg, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
    const int32_t* __restrict__ lab, int32_t* tkeys, uint32_t* tcnts) {
    __shared__ int32_t s_keys[kHeavySlots];
    __shared__ uint32_t s_cnts[kHeavySlots];
    const int32_t hv = chunk_vtx[blockIdx.x];
    const int32_t v = list[hv];
    const int32_t beg = chunk_beg[blockIdx.x], end = min(beg + kHeavyChunk, rowptr[v + 1]);
    for (int j = threadIdx.x; j < kHeavySlots; j += kBT) {
        s_keys[j] = kEmpty;
        s_cnts[j] = 0;
    }
    __syncthreads();
    for (int32_t j = beg + threadIdx.x; j < end; j += kBT) {
        const int32_t x = lab[col[j]];
        uint32_t h = hash_slot(x, kHeavySlots - 1);
        for (;;) {
            const int32_t prev = atomicCAS(&s_keys[h], kEmpty, x);
            if (prev == kEmpty || prev == x) {
                atomicAdd(&s_cnts[h], 1u);
                break;
            }
            h = (h + 1) & (kHeavySlots - 1);
        }
    }
    __syncthreads();
    const int64_t b0 = base[hv];
    const uint32_t mask = static_cast<uint32_t>(base[hv + 1] - b0 - 1);
    int32_t* keys = tkeys + b0;
    uint32_t* cnts = tcnts + b0;
    for (int j = threadIdx.x; j < kHeavySlots; j += kBT) {
        const int32_t x = s_keys[j];
        if (x == kEmpty) continue;
        uint32_t h = hash_slot(x, mask);
        for (;;) {
            const int32_t prev = atomicCAS(&keys[h], kEmpty, x);
            if (prev == kEmpty || prev == x) {
                atomicAdd(&cnts[h], s_cnts[j]);
                break;
            }
            h = (h + 1) & mask;
        }
    }
}

__global__ void __launch_bounds__(kBT) k_cdlp_heavy_pick(const int32_t* __restrict__ list,
                                                         const int64_t* __restrict__ base, int32_t* tkeys,
                                                         uint32_t* tcnts, int32_t* __restrict__ out) {
    __shared__ uint64_t s_red[kBT / 64];
    const int32_t hv = blockIdx.x;
    const int64_t b0 = base[hv], size = base[hv + 1] - b0;
    uint64_t best = 0;
    for (int64_t j = threadIdx.x; j < size; j += kBT) {
        const int32_t k = tkeys[b0 + j];
        if (k != kEmpty) {
            const uint64_t key = pack(tcnts[b0 + j], k);
            best = key > best ? key : best;
            tkeys[b0 + j] = kEmpty;
            tcnts[b0 + j] = 0;
        }
    }
    best = wave_max(best);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (int w = 0; w < kBT / 64; ++w) t = s_red[w] > t ? s_red[w] : t;
        out[list[hv]] = unpack_label(t);
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

// One workgroup per heavy vertex; the in-list in 256-element steps.
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

// Degree class of a vertex (0 lane, 1 wave, 2 LDS hash, 3 heavy; 4: no
// neighbours, keeps its label) as the key of a stable sort by class.
__global__ void k_cdlp_class(const int32_t* __restrict__ rp, int32_t n, uint32_t* __restrict__ key,
                             int32_t* __restrict__ val) {
    for (int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < n;
         v += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int32_t d = rp[v + 1] - rp[v];
        key[v] = d == 0 ? 4u : d <= kSmall ? 0u : d <= kWave ? 1u : d <= kLds ? 2u : 3u;
        val[v] = static_cast<int32_t>(v);
    }
}

// off[c] = first position of class c in the sorted keys; off[5] = n.
__global__ void k_cdlp_class_offsets(const uint32_t* __restrict__ key, int32_t n, int32_t* __restrict__ off) {
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e <= n;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t prev = e > 0 ? static_cast<int64_t>(key[e - 1]) : -1;
        const int64_t cur = e < n ? static_cast<int64_t>(key[e]) : 5;
        for (int64_t k = prev + 1; k <= cur; ++k) off[k] = static_cast<int32_t>(e);
    }
}

// Per heavy vertex: hash slots (the power of two >= 2d) and chunk count.
__global__ void k_cdlp_heavy_sizes(const int32_t* __restrict__ rp, const int32_t* __restrict__ heavy, int32_t h,
                                   int64_t* __restrict__ cap, int32_t* __restrict__ nch) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < h;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int32_t v = heavy[i], d = rp[v + 1] - rp[v];
        int64_t c = 1;
        while (c < 2LL * d) c <<= 1;
        cap[i] = c;
        nch[i] = (d + kHeavyChunk - 1) / kHeavyChunk;
    }
}

__global__ void k_cdlp_heavy_chunks(const int32_t* __restrict__ rp, const int32_t* __restrict__ heavy, int32_t h,
                                    const int32_t* __restrict__ choff, int32_t* __restrict__ cv,
                                    int32_t* __restrict__ cb) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < h;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int32_t v = heavy[i];
        int32_t o = choff[i];
        for (int32_t e = rp[v]; e < rp[v + 1]; e += kHeavyChunk, ++o) {
            cv[o] = static_cast<int32_t>(i);
            cb[o] = e;
        }
    }
}

// The class lists (ids ascending within a class) from a stable radix sort
// of the class keys, then the heavy vertices' hash-table offsets and
// 4096-edge chunks from scans; all on the GPU.
int build_cdlp_layout(ghet_graph* g, hipStream_t s) {
    CDLPLayout& L = g->cdlp;
    if (L.built) return GHET_OK;
    const int32_t n = g->n;
    const int32_t* rp = g->all.rowptr;
    uint32_t *ka = nullptr, *kb = nullptr;
    int32_t *va = nullptr, *vb = nullptr, *off = nullptr, *nch = nullptr, *choff = nullptr, *aux32 = nullptr;
    int64_t *cap = nullptr, *aux64 = nullptr, *hbase = nullptr;
    void* tmp = nullptr;
    size_t tb = 0;
    hipError_t e = hipSuccess;
    auto get = [&](void** q, size_t b) {
        if (e == hipSuccess) e = hipMalloc(q, std::max<size_t>(b, 256));
    };
    get(reinterpret_cast<void**>(&ka), 4 * size_t(n));
    get(reinterpret_cast<void**>(&kb), 4 * size_t(n));
    get(reinterpret_cast<void**>(&va), 4 * size_t(n));
    get(reinterpret_cast<void**>(&vb), 4 * size_t(n));
    get(reinterpret_cast<void**>(&off), 4 * 8);
    const unsigned gn = grid_cap(ceil_div(n, 256));
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_cdlp_class, dim3(gn), dim3(256), 0, s, rp, n, ka, va);
        e = hipGetLastError();
    }
    rocprim::double_buffer<uint32_t> kd(ka, kb);
    rocprim::double_buffer<int32_t> vd(va, vb);
    if (e == hipSuccess) e = rocprim::radix_sort_pairs(nullptr, tb, kd, vd, static_cast<unsigned>(n), 0, 3, s);
    get(&tmp, tb);
    if (e == hipSuccess) e = rocprim::radix_sort_pairs(tmp, tb, kd, vd, static_cast<unsigned>(n), 0, 3, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_cdlp_class_offsets, dim3(grid_cap(ceil_div(int64_t(n) + 1, 256))), dim3(256), 0, s,
                           kd.current(), n, off);
        e = hipGetLastError();
    }
    int32_t h_off[6] = {};
    if (e == hipSuccess) e = hipMemcpyAsync(h_off, off, sizeof(h_off), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    const int32_t* sorted = vd.current();
    const int32_t nh = h_off[4] - h_off[3];
    const int32_t* heavy = sorted + h_off[3];
    get(reinterpret_cast<void**>(&cap), 8 * size_t(nh));
    get(reinterpret_cast<void**>(&hbase), 8 * (size_t(nh) + 1));
    get(reinterpret_cast<void**>(&nch), 4 * size_t(nh));
    get(reinterpret_cast<void**>(&choff), 4 * (size_t(nh) + 1));
    get(reinterpret_cast<void**>(&aux32), 4 * size_t(scan_aux_ints(nh)));
    get(reinterpret_cast<void**>(&aux64), 8 * size_t(scan_aux_ints(nh)));
    if (e == hipSuccess && nh > 0) {
        hipLaunchKernelGGL(k_cdlp_heavy_sizes, dim3(grid_cap(ceil_div(nh, 256))), dim3(256), 0, s, rp, heavy, nh, cap, nch);
        e = hipGetLastError();
    }
    if (e == hipSuccess && exclusive_scan_i64(cap, hbase, nh, aux64, s) != GHET_OK) e = hipErrorUnknown;
    if (e == hipSuccess && exclusive_scan_i32(nch, choff, nh, aux32, s) != GHET_OK) e = hipErrorUnknown;
    int64_t slots = 0;
    int32_t chunks = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&slots, hbase + nh, 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(&chunks, choff + nh, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    size_t bytes = 0;
    for (int c = 0; c < 4; ++c) bytes += (4 * size_t(h_off[c + 1] - h_off[c]) + 255) & ~size_t(255);
    bytes += ((8 * (size_t(nh) + 1) + 255) & ~size_t(255)) + 2 * ((4 * size_t(chunks) + 255) & ~size_t(255));
    bytes += 2 * ((4 * size_t(slots) + 255) & ~size_t(255));
    if (e == hipSuccess) e = hipMalloc(&L.mem, std::max<size_t>(bytes, 256));
    if (e == hipSuccess) {
        char* p = static_cast<char*>(L.mem);
        auto take = [&](size_t b) { char* r = p; p += (b + 255) & ~size_t(255); return r; };
        for (int c = 0; c < 4 && e == hipSuccess; ++c) {
            L.count[c] = h_off[c + 1] - h_off[c];
            L.list[c] = reinterpret_cast<int32_t*>(take(4 * size_t(L.count[c])));
            if (L.count[c] > 0)
                e = hipMemcpyAsync(L.list[c], sorted + h_off[c], 4 * size_t(L.count[c]), hipMemcpyDeviceToDevice, s);
        }
        L.heavy_base = reinterpret_cast<int64_t*>(take(8 * (size_t(nh) + 1)));
        if (e == hipSuccess) e = hipMemcpyAsync(L.heavy_base, hbase, 8 * (size_t(nh) + 1), hipMemcpyDeviceToDevice, s);
        L.heavy_chunks = chunks;
        L.heavy_chunk_vtx = reinterpret_cast<int32_t*>(take(4 * size_t(chunks)));
        L.heavy_chunk_beg = reinterpret_cast<int32_t*>(take(4 * size_t(chunks)));
        if (e == hipSuccess && nh > 0) {
            hipLaunchKernelGGL(k_cdlp_heavy_chunks, dim3(grid_cap(ceil_div(nh, 256))), dim3(256), 0, s, rp, heavy, nh,
                               choff, L.heavy_chunk_vtx, L.heavy_chunk_beg);
            e = hipGetLastError();
        }
        L.heavy_slots = slots;
        L.heavy_keys = reinterpret_cast<int32_t*>(take(4 * size_t(slots)));
        L.heavy_cnts = reinterpret_cast<uint32_t*>(take(4 * size_t(slots)));
        if (e == hipSuccess && slots) e = hipMemsetAsync(L.heavy_keys, 0xFF, 4 * size_t(slots), s);
        if (e == hipSuccess && slots) e = hipMemsetAsync(L.heavy_cnts, 0, 4 * size_t(slots), s);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    for (void* q : {static_cast<void*>(ka), static_cast<void*>(kb), static_cast<void*>(va), static_cast<void*>(vb),
                    static_cast<void*>(off), static_cast<void*>(cap), static_cast<void*>(hbase), static_cast<void*>(nch),
                    static_cast<void*>(choff), static_cast<void*>(aux32), static_cast<void*>(aux64), tmp})
        if (q) hipFree(q);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        if (L.mem) hipFree(L.mem);
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
    int32_t *la, *lb;
    GHET_CHECK(g->ws.get(WS_LABEL_A, 4 * size_t(n), &p)); la = static_cast<int32_t*>(p);
    GHET_CHECK(g->ws.get(WS_LABEL_B, 4 * size_t(n), &p)); lb = static_cast<int32_t*>(p);
    const bool timing = st != nullptr;
    hipEvent_t* kev = nullptr;
    if (timing) GHET_CHECK(kernel_events(g, 2 * std::max(iters, 1), &kev));

    GHET_HIP(hipEventRecord(g->ev_begin, s));
    hipLaunchKernelGGL(k_label_init, dim3(grid_cap(ceil_div(n, 256))), dim3(256), 0, s, la, lb, n);
    // the identity-label first iteration has a closed form (see k_cdlp_first_*);
    // it needs the in-degrees to split each all-list, or an undirected graph
    const bool first_closed = !g->directed || (g->in.present() && g->in.rowptr != g->all.rowptr);
    const int32_t* in_rp = g->directed ? g->in.rowptr : nullptr;
    for (int32_t it = 0; it < iters; ++it) {
        if (timing) GHET_HIP(hipEventRecord(kev[2 * it], s));
        const int32_t* A = g->all.rowptr;
        const int32_t* C = g->all.col;
        if (it == 0 && first_closed) {
            // vertices with no neighbours are in no list and keep their id (lb holds ids)
            if (L.count[0])
                hipLaunchKernelGGL(k_cdlp_first_small, dim3(grid_cap(ceil_div(L.count[0], kBT))), dim3(kBT), 0, s,
                                   L.list[0], L.count[0], A, C, in_rp, lb);
            for (int c = 1; c <= 2; ++c)
                if (L.count[c])
                    hipLaunchKernelGGL(k_cdlp_first_wave, dim3(grid_cap(ceil_div(L.count[c], kBT / 64))), dim3(kBT), 0,
                                       s, L.list[c], L.count[c], A, C, in_rp, lb);
            if (L.count[3])
                hipLaunchKernelGGL(k_cdlp_first_block, dim3(grid_cap(L.count[3])), dim3(kBT), 0, s, L.list[3],
                                   L.count[3], A, C, in_rp, lb);
            if (timing) GHET_HIP(hipEventRecord(kev[2 * it + 1], s));
            std::swap(la, lb);
            continue;
        }
        if (L.count[0])
            hipLaunchKernelGGL(k_cdlp_small, dim3(grid_cap(ceil_div(L.count[0], kBT))), dim3(kBT), 0, s, L.list[0],
                               L.count[0], A, C, la, lb);
        if (L.count[1])
            hipLaunchKernelGGL(k_cdlp_wave, dim3(grid_cap(ceil_div(L.count[1], kBT / 64))), dim3(kBT), 0, s,
                               L.list[1], L.count[1], A, C, la, lb);
        if (L.count[2])
            hipLaunchKernelGGL(k_cdlp_lds, dim3(grid_cap(L.count[2])), dim3(kBT), 0, s, L.list[2], L.count[2], A, C,
                               la, lb);
        if (L.count[3]) {
            hipLaunchKernelGGL(k_cdlp_heavy_count, dim3(L.heavy_chunks), dim3(kBT), 0, s, L.list[3], L.heavy_base,
                               L.heavy_chunk_vtx, L.heavy_chunk_beg, A, C, la, L.heavy_keys, L.heavy_cnts);
            hipLaunchKernelGGL(k_cdlp_heavy_pick, dim3(L.count[3]), dim3(kBT), 0, s, L.list[3], L.heavy_base,
                               L.heavy_keys, L.heavy_cnts, lb);
        }
        if (timing) GHET_HIP(hipEventRecord(kev[2 * it + 1], s));
        std::swap(la, lb);
    }
    GHET_HIP(hipEventRecord(g->ev_end, s));
    GHET_HIP(hipGetLastError());
    if (label_out) GHET_CHECK(copy_out(g, label_out, la, 4 * size_t(n), s));
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
        for (int c = 0; c < 4; ++c) st->extra[c] = L.count[c];
    }
    return GHET_OK;
}

}  // namespace ghet
