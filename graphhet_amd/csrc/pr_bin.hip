// PageRank by destination bins on gfx950 (propagation blocking over a static
// layout).
//
// Reference: GPU_PR (GPU_PR_csr.cuh:34-117) gathers pr[u] / outdeg(u) per
// in-edge, one thread per vertex; semantics as in pr_common.hpp
// (CPU_PageRank.hpp:12-56).
//
// The pull gather (pr.hip) reads contrib[src] at random: one L2 request per
// edge, and a 64-128 B line per 4 useful bytes for the sources that miss L2.
// This form moves every random access into LDS and leaves global memory with
// streams only.  The in-edges are grouped once per graph by (source window,
// destination bin):
//   * a source window is 2^15 consecutive ids: its FP32 contributions are
//     128 KB, one LDS buffer;
//   * a destination bin is a run of consecutive ids, at most 2^14 of them
//     (128 KB of 64-bit accumulators) and about kBinEdges in-edges.
// One iteration is two launches:
//   1. k_prb_gather: per item (an edge range of one window) the window's
//      contributions are copied to LDS, and each edge's contribution is read
//      from there and written to val[e] in layout order.  2 B read + 4 B
//      written per edge, both sequential.
//   2. k_prb_accum: per bin, the bin's segment of every window is read (4 B
//      value + 2 B destination per edge, runs of consecutive edges); each
//      value is converted to fixed point with 62 fraction bits and added to
//      its destination's LDS accumulator with an integer atomic.  Then every
//      vertex of the bin is finished: rank, next contribution, sink mass and,
//      on the last iteration, the FP64 rank.
// Integer addition is exact and order-free, so a vertex's sum is the exact
// sum of its FP32 contributions truncated below 2^-62 per term, whatever the
// schedule: runs are bit-reproducible and need no compensation.  A row sum
// is at most the total rank (1), far below the 2^64 / 2^62 = 4 the
// accumulator holds.
#include <algorithm>
#include <numeric>

#include "graph_internal.hpp"
#include "pr_common.hpp"

#include <rocprim/rocprim.hpp>

namespace ghet {
namespace {

constexpr int kWinShift = 15;
constexpr int kWin = 1 << kWinShift;      // sources per window (128 KB of FP32 in LDS)
constexpr int kBinMax = 1 << 14;          // destinations per bin (128 KB of u64 in LDS)
constexpr int kMaxWin = 2048;             // windows the accumulate kernel's LDS tables hold
constexpr int64_t kBinEdges = 1 << 17;    // in-edges a bin aims for
constexpr int64_t kItemEdges = 1 << 17;   // edges per gather item
constexpr int kPB = 1024;                 // threads per block, both kernels (one block per CU)
constexpr int kQ = 8;                     // edges per thread per step
constexpr int kChunk = 64 * kQ;           // edges per wave per step (accumulate)
static_assert(kMaxWin <= 2 * kPB, "the segment scan gives each thread two windows");

// x >= 0 in FP32 -> x * 2^62 as an unsigned integer, truncated.  x =
// mant * 2^(ex - 150) for a normal x, so x * 2^62 = mant * 2^(ex - 88).
// Contributions are at most 1 (ex <= 127), so the result stays below 2^63.
__device__ __forceinline__ unsigned long long to_fix62(float x) {
    const uint32_t bits = __float_as_uint(x);
    const int ex = static_cast<int>(bits >> 23);
    if (ex == 0) return 0ull;  // zero or subnormal: below 2^-126, under the resolution
    const unsigned long long mant = (bits & 0x7FFFFFu) | 0x800000u;
    const int sh = ex - 88;
    return sh >= 0 ? (mant << sh) : (sh > -24 ? (mant >> -sh) : 0ull);
}

__global__ void __launch_bounds__(kPB) k_prb_gather(const int4* __restrict__ items, const uint16_t* __restrict__ src,
                                                    const float* __restrict__ contrib, int32_t n,
                                                    float* __restrict__ val) {
    __shared__ float s_win[kWin];
    const int4 it = items[blockIdx.x];
    const int64_t w0 = static_cast<int64_t>(it.x) << kWinShift;
    const int32_t wn = static_cast<int32_t>(n - w0 < kWin ? n - w0 : kWin);
    // contrib + w0 is 16-byte aligned: w0 is a multiple of 2^15
    const int32_t wn4 = wn >> 2;
    const float4* __restrict__ c4 = reinterpret_cast<const float4*>(contrib + w0);
    float4* s4 = reinterpret_cast<float4*>(s_win);
    for (int32_t i = threadIdx.x; i < wn4; i += kPB) s4[i] = c4[i];
    for (int32_t i = (wn4 << 2) + threadIdx.x; i < wn; i += kPB) s_win[i] = contrib[w0 + i];
    __syncthreads();
    const int32_t e1 = it.z;
    for (int32_t b = it.y + static_cast<int32_t>(threadIdx.x); b < e1; b += kPB * kQ) {
        uint32_t si[kQ];
#pragma unroll
        for (int j = 0; j < kQ; ++j) {
            const int32_t e = b + j * kPB;
            si[j] = e < e1 ? src[e] : 0u;
        }
#pragma unroll
        for (int j = 0; j < kQ; ++j) {
            const int32_t e = b + j * kPB;
            if (e < e1) val[e] = s_win[si[j]];
        }
    }
}

// Last segment w in [lo, hi) with s_pre[w] <= k, given s_pre[lo] <= k <
// s_pre[hi] (uniform k: every lane reads the same LDS words).
__device__ __forceinline__ int prb_seg_of(const int32_t* s_pre, int32_t k, int lo, int hi) {
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (__builtin_amdgcn_readfirstlane(s_pre[mid]) <= k) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Loads one wave chunk of a bin's edges (k in [c0, c0 + kChunk), lane j*64 +
// lane).  The chunk's first and last segments come from two wave-uniform
// searches; each lane then finds its edge's segment between them in a fixed
// number of steps, so no lane diverges, and the loads are issued together
// with no branch between them (lanes past T load edge 0 and are masked
// where the values are used).
__device__ __forceinline__ void prb_load_chunk(int32_t c0, int32_t T, int32_t num_win, const int32_t* s_pre,
                                               const int32_t* s_delta, const uint16_t* __restrict__ dst,
                                               const float* __restrict__ val, int lane, uint32_t (&dd)[kQ],
                                               float (&vv)[kQ]) {
    c0 = __builtin_amdgcn_readfirstlane(c0);
    const int32_t c1 = min(c0 + kChunk, T) - 1;
    const int w0 = prb_seg_of(s_pre, c0, 0, num_win);
    const int w1 = prb_seg_of(s_pre, c1, w0, num_win);
    const int steps = __builtin_amdgcn_readfirstlane(32 - __clz(w1 - w0));  // ceil(log2(w1 - w0 + 1))
    int32_t ei[kQ];
#pragma unroll
    for (int j = 0; j < kQ; ++j) {
        const int32_t k = min(c0 + j * 64 + lane, c1);
        int lo = w0, hi = w1 + 1;
        for (int t = 0; t < steps; ++t) {
            const int mid = (lo + hi) >> 1;
            const bool le = s_pre[mid] <= k;
            lo = le ? mid : lo;
            hi = le ? hi : mid;
        }
        ei[j] = c0 + j * 64 + lane <= c1 ? k + s_delta[lo] : 0;
    }
#pragma unroll
    for (int j = 0; j < kQ; ++j) {
        dd[j] = dst[ei[j]];
        vv[j] = val[ei[j]];
    }
}

// One block per bin (bins in descending edge order).  Each wave takes
// 512-edge chunks of the bin's edges, concatenated over the windows, and
// keeps two in flight: the next chunk's loads are issued before the current
// chunk's values are added.
__global__ void __launch_bounds__(kPB) k_prb_accum(
    const int32_t* __restrict__ bin_order, const int32_t* __restrict__ bin_start, const int2* __restrict__ seg,
    int32_t num_win, const uint16_t* __restrict__ dst, const float* __restrict__ val,
    const int32_t* __restrict__ outdeg, const PRScalars* __restrict__ scal, double nd, double damping,
    float* __restrict__ next, double* __restrict__ part, double* __restrict__ rank_out, int last) {
    __shared__ unsigned long long s_acc[kBinMax];
    __shared__ int32_t s_pre[kMaxWin + 1];
    __shared__ int32_t s_delta[kMaxWin];  // first edge of segment w minus s_pre[w]
    __shared__ int32_t s_wt[kPB / 64];
    __shared__ double s_red[kPB / 64];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int32_t b = bin_order[blockIdx.x];
    const int32_t v0 = bin_start[b], nv = bin_start[b + 1] - v0;
    for (int32_t i = tid; i < nv; i += kPB) s_acc[i] = 0ull;
    // exclusive scan of the segment lengths, two windows per thread
    {
        const int2* sb = seg + static_cast<int64_t>(b) * num_win;
        const int i0 = 2 * tid;
        const int2 a = i0 < num_win ? sb[i0] : make_int2(0, 0);
        const int2 c = i0 + 1 < num_win ? sb[i0 + 1] : make_int2(0, 0);
        const int32_t mine = a.y + c.y;
        int32_t incl = mine;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int32_t y = __shfl_up(incl, off);
            if (lane >= off) incl += y;
        }
        if (lane == 63) s_wt[wave] = incl;
        __syncthreads();
        int32_t wbase = 0;
        for (int w = 0; w < wave; ++w) wbase += s_wt[w];
        const int32_t excl = wbase + incl - mine;
        if (i0 < num_win) {
            s_pre[i0] = excl;
            s_delta[i0] = a.x - excl;
        }
        if (i0 + 1 < num_win) {
            s_pre[i0 + 1] = excl + a.y;
            s_delta[i0 + 1] = c.x - (excl + a.y);
        }
        if (tid == kPB - 1) s_pre[num_win] = wbase + incl;
        __syncthreads();
    }
    const int32_t T = s_pre[num_win];
    constexpr int32_t kStride = (kPB / 64) * kChunk;
    uint32_t dd[kQ];
    float vv[kQ];
    int32_t c0 = wave * kChunk;
    if (c0 < T) prb_load_chunk(c0, T, num_win, s_pre, s_delta, dst, val, lane, dd, vv);
    for (; c0 < T; c0 += kStride) {
        uint32_t dn[kQ];
        float vn[kQ];
        // the last pass reloads the current chunk (in bounds, never used)
        prb_load_chunk(c0 + kStride < T ? c0 + kStride : c0, T, num_win, s_pre, s_delta, dst, val, lane, dn, vn);
#pragma unroll
        for (int j = 0; j < kQ; ++j)
            if (c0 + j * 64 + lane < T) atomicAdd(&s_acc[dd[j] & (kBinMax - 1)], to_fix62(vv[j]));
#pragma unroll
        for (int j = 0; j < kQ; ++j) {
            dd[j] = dn[j];
            vv[j] = vn[j];
        }
    }
    __syncthreads();
    const double base = scal->sink * damping / nd + (1.0 - damping) / nd;
    double sink = 0;
    for (int32_t i = tid; i < nv; i += kPB) {
        const double pr = base + damping * (static_cast<double>(s_acc[i]) * 0x1p-62);
        const int32_t d = outdeg[v0 + i];
        if (d > 0) next[v0 + i] = static_cast<float>(pr / d);
        else sink += pr;
        if (last) rank_out[v0 + i] = pr;
    }
    const double t = block_sum_f64<kPB>(sink, s_red);
    if (tid == 0) part[b] = t;
}

// ---------------------------------------------------------------------------
// Layout construction (once per graph, on the GPU; bins and items on the host
// from the in-offsets and the window offsets).  The edges are read from the
// out-CSR; bins are sized from the in-offsets.

__global__ void k_prb_vbin(const int32_t* __restrict__ bin_start, int32_t num_bin, int32_t* __restrict__ vbin) {
    const int lane = lane_id();
    const int64_t wave = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const int64_t waves = (static_cast<int64_t>(gridDim.x) * blockDim.x) >> 6;
    for (int64_t b = wave; b < num_bin; b += waves)
        for (int32_t v = bin_start[b] + lane; v < bin_start[b + 1]; v += 64) vbin[v] = static_cast<int32_t>(b);
}

// Per edge u -> v: key = window(u) * num_bin + bin(v); word = (u within its
// window) | (v within its bin) << 16.  One wave per row of the out-CSR, so
// after the stable sort each segment's edges are ordered by source: the
// destinations of consecutive edges are then spread over the bin, and a
// wave's LDS atomics seldom meet on one accumulator (in destination order a
// vertex's edges from one window are adjacent and their atomics serialise).
__global__ void k_prb_keys(const int32_t* __restrict__ rp, const int32_t* __restrict__ col, int32_t n,
                           const int32_t* __restrict__ vbin, const int32_t* __restrict__ bin_start, uint32_t num_bin,
                           uint32_t* __restrict__ key, uint32_t* __restrict__ word) {
    const int lane = lane_id();
    const int64_t wave = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const int64_t waves = (static_cast<int64_t>(gridDim.x) * blockDim.x) >> 6;
    for (int64_t u = wave; u < n; u += waves) {
        const uint32_t kw = static_cast<uint32_t>(u >> kWinShift) * num_bin;
        const uint32_t sl = static_cast<uint32_t>(u & (kWin - 1));
        for (int32_t e = rp[u] + lane; e < rp[u + 1]; e += 64) {
            const int32_t v = col[e];
            const int32_t b = vbin[v];
            key[e] = kw + static_cast<uint32_t>(b);
            word[e] = (static_cast<uint32_t>(v - bin_start[b]) << 16) | sl;
        }
    }
}

// soff[k] = first position of key k in the sorted keys (an empty key gets
// the position of the next present one); soff[num_keys] = count.
__global__ void k_prb_soff(const uint32_t* __restrict__ key, int64_t count, uint32_t num_keys,
                           int32_t* __restrict__ soff) {
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e <= count;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t prev = e > 0 ? static_cast<int64_t>(key[e - 1]) : -1;
        const int64_t cur = e < count ? static_cast<int64_t>(key[e]) : static_cast<int64_t>(num_keys);
        for (int64_t k = prev + 1; k <= cur; ++k) soff[k] = static_cast<int32_t>(e);
    }
}

__global__ void k_prb_split(const uint32_t* __restrict__ word, int64_t m, uint16_t* __restrict__ src,
                            uint16_t* __restrict__ dst) {
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < m;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const uint32_t x = word[e];
        src[e] = static_cast<uint16_t>(x & 0xFFFFu);
        dst[e] = static_cast<uint16_t>(x >> 16);
    }
}

// Bin-major segment table, and the first edge of every window (win[w] =
// soff[w * num_bin], win[num_win] = m).
__global__ void k_prb_seg(const int32_t* __restrict__ soff, int32_t num_win, int32_t num_bin, int2* __restrict__ seg,
                          int32_t* __restrict__ win) {
    const int64_t total = static_cast<int64_t>(num_win) * num_bin;
    for (int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; q <= total;
         q += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        if (q < total) {
            const int64_t b = q / num_win, w = q - b * num_win;
            const int64_t k = w * num_bin + b;
            seg[q] = make_int2(soff[k], soff[k + 1] - soff[k]);
        }
        if (q % num_bin == 0 && q / num_bin <= num_win) win[q / num_bin] = soff[q];
    }
}

__global__ void k_prb_outdeg(const int32_t* __restrict__ out_rp, int32_t n, int32_t* __restrict__ outdeg) {
    for (int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < n;
         v += static_cast<int64_t>(gridDim.x) * blockDim.x)
        outdeg[v] = out_rp[v + 1] - out_rp[v];
}

unsigned grid_for(int64_t threads, int bt = 256) {
    return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(ceil_div(threads, bt), 65536)));
}

int build_prb_layout(ghet_graph* g, hipStream_t s) {
    PRBinLayout& L = g->prb;
    if (L.built) return GHET_OK;
    const int32_t n = g->n;
    const int64_t m = g->in.m;
    if (!g->out.present() || g->out.col == nullptr || g->out.m != m)
        return set_error(GHET_ERR_INVALID, "PageRank bins need the out-CSR with the in-CSR's edges");
    if (!pr_binned_fits(n, m)) return set_error(GHET_ERR_UNSUPPORTED, "PageRank bins: graph too large for the bin tables");
    // Bins: consecutive destinations, closed before a bin would pass
    // kBinMax vertices or kBinEdges in-edges (a vertex with more in-edges
    // than that is a bin of its own).
    const std::vector<int32_t>& rp = g->h_in_rowptr;
    std::vector<int32_t> bs{0};
    std::vector<int64_t> bedges;
    {
        int64_t e = 0;
        int32_t start = 0;
        for (int32_t v = 0; v < n; ++v) {
            const int64_t d = rp[v + 1] - rp[v];
            if (v > start && (v - start == kBinMax || e + d > kBinEdges)) {
                bs.push_back(v);
                bedges.push_back(e);
                start = v;
                e = 0;
            }
            e += d;
        }
        bs.push_back(n);
        bedges.push_back(e);
    }
    const int32_t NB = static_cast<int32_t>(bs.size() - 1);
    const int32_t NW = static_cast<int32_t>(ceil_div(n, kWin));
    const int64_t K = int64_t(NW) * NB;
    if (K >= INT32_MAX) return set_error(GHET_ERR_UNSUPPORTED, "PageRank bins: %lld segments", static_cast<long long>(K));
    std::vector<int32_t> order(NB);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return bedges[a] > bedges[b]; });
    const int64_t max_items = m / kItemEdges + NW + 1;

    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t me = std::max<size_t>(size_t(m), 1);
    const size_t bytes = 2 * al(2 * me) + al(4 * me) + al(8 * size_t(K)) + al(4 * (size_t(NB) + 1)) + al(4 * size_t(NB)) +
                         al(16 * size_t(max_items)) + al(4 * size_t(n));
    GHET_HIP(hipMalloc(&L.mem, bytes));
    {
        char* p = static_cast<char*>(L.mem);
        auto take = [&](size_t b) { char* r = p; p += al(b); return r; };
        L.src = reinterpret_cast<uint16_t*>(take(2 * me));
        L.dst = reinterpret_cast<uint16_t*>(take(2 * me));
        L.val = reinterpret_cast<float*>(take(4 * me));
        L.seg = reinterpret_cast<int2*>(take(8 * size_t(K)));
        L.bin_start = reinterpret_cast<int32_t*>(take(4 * (size_t(NB) + 1)));
        L.bin_order = reinterpret_cast<int32_t*>(take(4 * size_t(NB)));
        L.items = reinterpret_cast<int4*>(take(16 * size_t(max_items)));
        L.outdeg = reinterpret_cast<int32_t*>(take(4 * size_t(n)));
    }
    auto fail = [&](const char* what, hipError_t e) {
        (void)hipGetLastError();
        (void)hipStreamSynchronize(s);
        hipFree(L.mem);
        L = PRBinLayout{};
        return set_error(GHET_ERR_HIP, "PageRank bins (%s): %s", what, hipGetErrorString(e));
    };
    std::vector<int32_t> win(static_cast<size_t>(NW) + 1);
    {
        DevTemp t;
        auto ok = [&](hipError_t e) {
            if (t.err == hipSuccess) t.err = e;
            return t.err == hipSuccess;
        };
        int32_t* vbin = t.get<int32_t>(n);
        uint32_t* ka = t.get<uint32_t>(me);
        uint32_t* kb = t.get<uint32_t>(me);
        uint32_t* wa = t.get<uint32_t>(me);
        uint32_t* wb = t.get<uint32_t>(me);
        int32_t* soff = t.get<int32_t>(size_t(K) + 1);
        int32_t* d_win = t.get<int32_t>(size_t(NW) + 1);
        if (t.err != hipSuccess) return fail("alloc", t.err);
        ok(hipMemcpyAsync(L.bin_start, bs.data(), 4 * bs.size(), hipMemcpyHostToDevice, s));
        ok(hipMemcpyAsync(L.bin_order, order.data(), 4 * order.size(), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_prb_vbin, dim3(grid_for(int64_t(NB) * 64)), dim3(256), 0, s, L.bin_start, NB, vbin);
        hipLaunchKernelGGL(k_prb_outdeg, dim3(grid_for(n)), dim3(256), 0, s, g->out.rowptr, n, L.outdeg);
        hipLaunchKernelGGL(k_prb_keys, dim3(grid_for(int64_t(n) * 64)), dim3(256), 0, s, g->out.rowptr, g->out.col, n,
                           vbin, L.bin_start, static_cast<uint32_t>(NB), ka, wa);
        ok(hipGetLastError());
        unsigned bits = 1;
        while ((int64_t(1) << bits) < K) ++bits;
        rocprim::double_buffer<uint32_t> kd(ka, kb), wd(wa, wb);
        if (m > 0) {
            size_t need = 0;
            ok(rocprim::radix_sort_pairs(nullptr, need, kd, wd, static_cast<unsigned>(m), 0, bits, s));
            void* tmp = t.get<char>(need);
            ok(rocprim::radix_sort_pairs(tmp, need, kd, wd, static_cast<unsigned>(m), 0, bits, s));
        }
        hipLaunchKernelGGL(k_prb_soff, dim3(grid_for(m + 1)), dim3(256), 0, s, kd.current(), m, static_cast<uint32_t>(K),
                           soff);
        hipLaunchKernelGGL(k_prb_split, dim3(grid_for(m)), dim3(256), 0, s, wd.current(), m, L.src, L.dst);
        hipLaunchKernelGGL(k_prb_seg, dim3(grid_for(K + 1)), dim3(256), 0, s, soff, NW, NB, L.seg, d_win);
        ok(hipGetLastError());
        ok(hipMemcpyAsync(win.data(), d_win, 4 * win.size(), hipMemcpyDeviceToHost, s));
        ok(hipStreamSynchronize(s));
        if (t.err != hipSuccess) return fail("build", t.err);
    }
    // Gather items: each window's edges cut into ranges of at most
    // kItemEdges; every index is checked here once.
    std::vector<int4> items;
    bool good = win[0] == 0 && win[NW] == m;
    for (int32_t w = 0; good && w < NW; ++w) {
        good = win[w] <= win[w + 1];
        for (int64_t e0 = win[w]; good && e0 < win[w + 1]; e0 += kItemEdges)
            items.push_back(make_int4(w, static_cast<int32_t>(e0), static_cast<int32_t>(std::min<int64_t>(win[w + 1], e0 + kItemEdges)), 0));
    }
    if (!good || static_cast<int64_t>(items.size()) > max_items) {
        hipFree(L.mem);
        L = PRBinLayout{};
        return set_error(GHET_ERR_INVALID, "PageRank bins: window offsets failed their check");
    }
    if (!items.empty()) {
        const hipError_t e = hipMemcpyAsync(L.items, items.data(), 16 * items.size(), hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return fail("items", e);
    }
    const hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail("items", e);
    L.num_win = NW;
    L.num_bin = NB;
    L.num_items = static_cast<int32_t>(items.size());
    L.built = true;
    return GHET_OK;
}

}  // namespace

bool pr_binned_fits(int32_t n, int64_t m) {
    return n > 0 && ceil_div(n, kWin) <= kMaxWin && m < INT32_MAX;
}

bool pr_binned_default(int32_t n, int64_t m) {
    return false && n > (1 << 21) && pr_binned_fits(n, m);
}

int run_pr_binned(ghet_graph* g, int32_t iters, double damping, double* pr_out, bool normalise, const ghet_opts* o,
                  ghet_run_stats* st) {
    const int32_t n = g->n;
    hipStream_t s = pick_stream(g, o);
    GHET_CHECK(build_prb_layout(g, s));
    const PRBinLayout& L = g->prb;
    const int32_t init_blocks = static_cast<int32_t>(std::min<int64_t>(ceil_div(n, 256), 4096));
    void* p;
    float *ca, *cb;
    double *rank, *out, *part;
    PRScalars* scal;
    GHET_CHECK(g->ws.get(WS_PR_CONTRIB, 4 * size_t(n), &p)); ca = static_cast<float*>(p);
    GHET_CHECK(g->ws.get(WS_LABEL_A, 4 * size_t(n), &p)); cb = static_cast<float*>(p);
    GHET_CHECK(g->ws.get(WS_PR_RANK, 8 * size_t(n), &p)); rank = static_cast<double*>(p);
    GHET_CHECK(g->ws.get(WS_PR_OUT, 8 * size_t(n), &p)); out = static_cast<double*>(p);
    GHET_CHECK(g->ws.get(WS_PR_PART, 8 * size_t(std::max(L.num_bin, init_blocks) + 1), &p)); part = static_cast<double*>(p);
    GHET_CHECK(g->ws.get(WS_COUNTERS, 256, &p)); scal = static_cast<PRScalars*>(p);
    const double inv_n = 1.0 / n;
    hipEvent_t* kev = nullptr;
    if (st) GHET_CHECK(kernel_events(g, 3 * iters, &kev));

    GHET_HIP(hipEventRecord(g->ev_begin, s));
    hipLaunchKernelGGL(k_pr_init, dim3(init_blocks), dim3(256), 0, s, L.outdeg, n, inv_n, ca, part);
    hipLaunchKernelGGL(k_pr_reduce, dim3(1), dim3(1024), 0, s, part, init_blocks, &scal->sink, static_cast<int*>(nullptr));
    if (iters == 0) GHET_CHECK(fill<double>(rank, inv_n, n, s));
    for (int32_t it = 0; it < iters; ++it) {
        const int last = (it == iters - 1);
        if (st) GHET_HIP(hipEventRecord(kev[3 * it], s));
        if (L.num_items > 0)
            hipLaunchKernelGGL(k_prb_gather, dim3(L.num_items), dim3(kPB), 0, s, L.items, L.src, ca, n, L.val);
        if (st) GHET_HIP(hipEventRecord(kev[3 * it + 1], s));
        hipLaunchKernelGGL(k_prb_accum, dim3(L.num_bin), dim3(kPB), 0, s, L.bin_order, L.bin_start, L.seg, L.num_win,
                           L.dst, L.val, L.outdeg, scal, static_cast<double>(n), damping, cb, part, rank, last);
        if (st) GHET_HIP(hipEventRecord(kev[3 * it + 2], s));
        hipLaunchKernelGGL(k_pr_reduce, dim3(1), dim3(1024), 0, s, part, L.num_bin, &scal->sink, static_cast<int*>(nullptr));
        std::swap(ca, cb);
    }
    const int norm = (normalise && iters > 0) ? 1 : 0;
    if (norm) {
        hipLaunchKernelGGL(k_pr_sum, dim3(init_blocks), dim3(256), 0, s, rank, n, part);
        hipLaunchKernelGGL(k_pr_reduce, dim3(1), dim3(1024), 0, s, part, init_blocks, &scal->total, static_cast<int*>(nullptr));
    }
    hipLaunchKernelGGL(k_pr_unpermute, dim3(init_blocks), dim3(256), 0, s, rank, static_cast<const int32_t*>(nullptr), n,
                       scal, norm, out);
    GHET_HIP(hipEventRecord(g->ev_end, s));
    GHET_HIP(hipGetLastError());
    if (pr_out) GHET_CHECK(copy_out(g, pr_out, out, 8 * size_t(n), s));
    GHET_HIP(hipMemcpyAsync(g->pinned, scal, sizeof(PRScalars), hipMemcpyDeviceToHost, s));
    GHET_HIP(hipStreamSynchronize(s));
    if (st) {
        double iter_ms = 0, gather_ms = 0;
        for (int32_t it = 0; it < iters; ++it) {
            float kms = 0, gms = 0;
            GHET_HIP(hipEventElapsedTime(&kms, kev[3 * it], kev[3 * it + 2]));
            GHET_HIP(hipEventElapsedTime(&gms, kev[3 * it], kev[3 * it + 1]));
            iter_ms += kms;
            gather_ms += gms;
        }
        float ms = 0;
        GHET_HIP(hipEventElapsedTime(&ms, g->ev_begin, g->ev_end));
        st->device_ms = ms;
        st->hot_kernel_ms = iter_ms;  // both launches of every iteration
        st->hot_kernel_launches = iters;
        st->levels = iters;
        const PRScalars* hs = reinterpret_cast<const PRScalars*>(g->pinned);
        st->extra[0] = hs->sink;
        st->extra[1] = hs->total;
        st->extra[2] = gather_ms;  // of which the window gathers (k_prb_gather)
        st->extra[3] = 2;          // form: 2 = destination bins
    }
    return GHET_OK;
}

}  // namespace ghet
