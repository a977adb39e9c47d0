// Pieces shared by the two PageRank forms (pr.hip: pull gather over a
// degree-relabeled in-CSR; pr_bin.hip: destination bins over a source-window
// layout): the scalars carried between iterations, the initial contributions,
// the ordered FP64 reductions and the final write of the ranks.
//
// Semantics (CPU_PageRank.hpp:12-56): pr0 = 1/N; each iteration
//     pr'[v] = (S*d/N + (1-d)/N) + d * sum_{u in in(v)} pr[u]/outdeg(u),
// S = sum of pr over vertices without out-edges.  Every reduction here runs
// in a fixed order, so a run is bit-reproducible.
#pragma once

#include "graph_internal.hpp"

namespace ghet {
namespace {

struct PRScalars {
    double sink;   // S of the current ranks
    double total;  // sum of ranks (normalisation)
};

// Block-wide FP64 sum in a fixed order (wave trees, then waves in index
// order); the result is valid in thread 0.  BT = threads per block.
template <int BT = 256>
__device__ __forceinline__ double block_sum_f64(double x, double* s_red) {
    x = wave_sum(x);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) s_red[w] = x;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < BT / 64; ++i) t += s_red[i];
    __syncthreads();
    return t;
}

// contrib[v] = pr0 / outdeg(v) in FP32 (0 for sinks); part[block] = the
// block's sink mass.
__global__ void __launch_bounds__(256) k_pr_init(const int32_t* __restrict__ outdeg, int32_t n, double pr0,
                                                 float* __restrict__ contrib, double* __restrict__ part) {
    __shared__ double s_red[256 / 64];
    double sink = 0;
    for (int64_t v = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; v < n;
         v += static_cast<int64_t>(gridDim.x) * 256) {
        const int32_t d = outdeg[v];
        if (d > 0) contrib[v] = static_cast<float>(pr0 / d);
        else {
            contrib[v] = 0.f;
            sink += pr0;
        }
    }
    const double t = block_sum_f64<256>(sink, s_red);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// Sums part[0..count) in index order (strided lanes, then fixed tree) and
// stores it in *out; also zeroes *zero when given.
__global__ void __launch_bounds__(1024) k_pr_reduce(const double* __restrict__ part, int32_t count, double* out,
                                                    int* zero) {
    __shared__ double s[1024];
    if (zero && threadIdx.x == 0) *zero = 0;  // a tile counter for the next launch
    double x = 0;
    for (int i = threadIdx.x; i < count; i += 1024) x += part[i];
    s[threadIdx.x] = x;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0];
}

__global__ void __launch_bounds__(256) k_pr_sum(const double* __restrict__ rank, int32_t n, double* __restrict__ part) {
    __shared__ double s_red[256 / 64];
    double x = 0;
    for (int64_t v = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; v < n;
         v += static_cast<int64_t>(gridDim.x) * 256)
        x += rank[v];
    const double t = block_sum_f64<256>(x, s_red);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// out[perm[r]] = rank[r] / total (total = 1 without normalisation); a null
// perm is the identity.
__global__ void k_pr_unpermute(const double* __restrict__ rank, const int32_t* __restrict__ perm, int32_t n,
                               const PRScalars* scal, int normalise, double* __restrict__ out) {
    const double inv = normalise ? 1.0 / scal->total : 1.0;
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; r < n;
         r += static_cast<int64_t>(gridDim.x) * blockDim.x)
        out[perm ? perm[r] : r] = normalise ? rank[r] * inv : rank[r];
}

}  // namespace
}  // namespace ghet
