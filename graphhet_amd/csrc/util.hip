// Device-wide primitives shared by the algorithms: fill, exclusive scan
// (reduce / scan-of-partials / apply, all deterministic) and reached-vertex
// statistics.
#include "graph_internal.hpp"

namespace ghet {

template <typename T>
__global__ void k_fill(T* __restrict__ p, T v, int64_t n) {
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

template <typename T>
int fill(T* p, T v, int64_t n, hipStream_t s) {
    if (n <= 0) return GHET_OK;
    int64_t blocks = std::min<int64_t>(ceil_div(n, 256), 8192);
    hipLaunchKernelGGL(k_fill<T>, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, p, v, n);
    GHET_HIP(hipGetLastError());
    return GHET_OK;
}
template int fill<int32_t>(int32_t*, int32_t, int64_t, hipStream_t);
template int fill<uint32_t>(uint32_t*, uint32_t, int64_t, hipStream_t);
template int fill<int64_t>(int64_t*, int64_t, int64_t, hipStream_t);
template int fill<uint64_t>(uint64_t*, uint64_t, int64_t, hipStream_t);
template int fill<double>(double*, double, int64_t, hipStream_t);
template int fill<float>(float*, float, int64_t, hipStream_t);

// ---------------------------------------------------------------- scan ----
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;  // 4096 elements per block

// Block-wide exclusive scan of one value per thread; returns the block total.
template <typename T>
__device__ T block_exclusive_scan(T x, T* lds_waves, T& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    T incl = x;
    for (int off = 1; off < 64; off <<= 1) {
        T y = __shfl_up(incl, off);
        if (lane >= off) incl += y;
    }
    if (lane == 63) lds_waves[wave] = incl;
    __syncthreads();
    T wave_off = 0, tot = 0;
    for (int w = 0; w < nw; ++w) {
        T s = lds_waves[w];
        if (w < wave) wave_off += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return wave_off + incl - x;
}

template <typename T>
__global__ void __launch_bounds__(kScanThreads) k_scan_reduce(const T* __restrict__ in, int64_t n, T* __restrict__ aux) {
    __shared__ T lds[kScanThreads / 64];
    int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
    T s = 0;
    for (int j = 0; j < kScanItems; ++j) {
        int64_t i = base + j * kScanThreads + threadIdx.x;
        if (i < n) s += in[i];
    }
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        T t = 0;
        for (int w = 0; w < kScanThreads / 64; ++w) t += lds[w];
        aux[blockIdx.x] = t;
    }
}

template <typename T>
__global__ void __launch_bounds__(1024) k_scan_aux(T* __restrict__ aux, int64_t nb) {
    __shared__ T lds[16];
    __shared__ T carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int64_t b = 0; b < nb; b += 1024) {
        int64_t i = b + threadIdx.x;
        T x = i < nb ? aux[i] : 0;
        T tot;
        T ex = block_exclusive_scan(x, lds, tot);
        if (i < nb) aux[i] = carry + ex;
        __syncthreads();
        if (threadIdx.x == 0) carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) aux[nb] = carry;
}

template <typename T>
__global__ void __launch_bounds__(kScanThreads) k_scan_apply(const T* in, T* out, int64_t n, const T* __restrict__ aux,
                                                              int64_t nb) {
    __shared__ T lds[kScanThreads / 64];
    int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile + static_cast<int64_t>(threadIdx.x) * kScanItems;
    T v[kScanItems];
    T s = 0;
    for (int j = 0; j < kScanItems; ++j) {
        int64_t i = base + j;
        v[j] = i < n ? in[i] : 0;
        s += v[j];
    }
    T tot;
    T run = block_exclusive_scan(s, lds, tot) + aux[blockIdx.x];
    for (int j = 0; j < kScanItems; ++j) {
        int64_t i = base + j;
        if (i < n) out[i] = run;
        run += v[j];
    }
    if (blockIdx.x == nb - 1 && threadIdx.x == 0) out[n] = aux[nb];
}

int64_t scan_aux_ints(int64_t n) { return ceil_div(std::max<int64_t>(n, 1), kScanTile) + 1; }

template <typename T>
int exclusive_scan(const T* in, T* out, int64_t n, T* aux, hipStream_t s) {
    if (n <= 0) {
        GHET_CHECK(fill<T>(out, 0, 1, s));
        return GHET_OK;
    }
    int64_t nb = ceil_div(n, kScanTile);
    hipLaunchKernelGGL(k_scan_reduce<T>, dim3(static_cast<unsigned>(nb)), dim3(kScanThreads), 0, s, in, n, aux);
    hipLaunchKernelGGL(k_scan_aux<T>, dim3(1), dim3(1024), 0, s, aux, nb);
    hipLaunchKernelGGL(k_scan_apply<T>, dim3(static_cast<unsigned>(nb)), dim3(kScanThreads), 0, s, in, out, n, aux, nb);
    GHET_HIP(hipGetLastError());
    return GHET_OK;
}

int exclusive_scan_i32(const int32_t* in, int32_t* out, int64_t n, int32_t* aux, hipStream_t s) {
    return exclusive_scan<int32_t>(in, out, n, aux, s);
}
int exclusive_scan_i64(const int64_t* in, int64_t* out, int64_t n, int64_t* aux, hipStream_t s) {
    return exclusive_scan<int64_t>(in, out, n, aux, s);
}

// ------------------------------------------------------- reached stats ----
// Counts n_r and m_r into slots (field 0 and 1); the host sums the slots.
template <typename Pred>
__device__ void reached_body(Pred reached, const int32_t* __restrict__ rowptr, int32_t n, CtrSlot* slots) {
    __shared__ unsigned long long s_red[4];
    unsigned long long cnt = 0, edges = 0;
    for (int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < n;
         v += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        if (reached(v)) {
            ++cnt;
            edges += rowptr[v + 1] - rowptr[v];
        }
    }
    cnt = block_sum_u64<256>(cnt, s_red);
    edges = block_sum_u64<256>(edges, s_red);
    if (threadIdx.x == 0) {
        slot_add(slots, 0, cnt);
        slot_add(slots, 1, edges);
    }
}

__global__ void __launch_bounds__(256) k_reached_i32(const int32_t* __restrict__ depth, const int32_t* __restrict__ rowptr,
                                                     int32_t n, CtrSlot* slots) {
    reached_body([&](int64_t v) { return depth[v] != kIntInf; }, rowptr, n, slots);
}
__global__ void __launch_bounds__(256) k_reached_f64(const double* __restrict__ dist, double inf,
                                                     const int32_t* __restrict__ rowptr, int32_t n, CtrSlot* slots) {
    reached_body([&](int64_t v) { return dist[v] != inf; }, rowptr, n, slots);
}

int reached_stats_i32(const int32_t* depth, const int32_t* rowptr, int32_t n, CtrSlot* slots, hipStream_t s) {
    GHET_CHECK(fill<uint64_t>(reinterpret_cast<uint64_t*>(slots), 0, kSlots * 8, s));
    unsigned blocks = static_cast<unsigned>(std::min<int64_t>(ceil_div(n, 256), 2048));
    hipLaunchKernelGGL(k_reached_i32, dim3(std::max(blocks, 1u)), dim3(256), 0, s, depth, rowptr, n, slots);
    GHET_HIP(hipGetLastError());
    return GHET_OK;
}
int reached_stats_f64(const double* dist, double inf, const int32_t* rowptr, int32_t n, CtrSlot* slots,
                      hipStream_t s) {
    GHET_CHECK(fill<uint64_t>(reinterpret_cast<uint64_t*>(slots), 0, kSlots * 8, s));
    unsigned blocks = static_cast<unsigned>(std::min<int64_t>(ceil_div(n, 256), 2048));
    hipLaunchKernelGGL(k_reached_f64, dim3(std::max(blocks, 1u)), dim3(256), 0, s, dist, inf, rowptr, n, slots);
    GHET_HIP(hipGetLastError());
    return GHET_OK;
}

}  // namespace ghet
