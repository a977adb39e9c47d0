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
// streams.  Once per graph the edges are grouped by (source window,
// destination bin):
//   * a source window is 2^15 consecutive ids: its contribution words are
//     128 KB, one LDS buffer;
//   * a destination bin is a run of consecutive ids, at most 2^14 of them
//     (128 KB of 64-bit accumulators) and about kBinEdges in-edges.
// The edges of window w going to bin b form segment (w, b).  Sources are
// stored window-major (segments of a window in bin order, each ordered by
// source); destinations and the per-iteration values bin-major (segments of
// a bin in window order).  One iteration is two launches:
//   1. k_prb_gather: per item (an edge range of one window) the window's
//      contribution words are copied to LDS with the item's segment table;
//      each edge reads its source's word from LDS and stores it at its
//      bin-major position (a segment is one contiguous run).  2 B read and
//      4 B written per edge.
//   2. k_prb_accum: per bin, the bin's values and destinations are read
//      sequentially (6 B per edge) and each value is added to its
//      destination's LDS accumulator with a 64-bit integer atomic; then every
//      vertex of the bin is finished: rank, next contribution word, sink mass
//      and, on the last iteration, the FP64 rank.
// Contribution words.  contrib = (float)(pr / outdeg) as in the gather form,
// stored as (s << 24) | m with value m * 2^(s - 62): the FP32 value in fixed
// point with 62 fraction bits (m is the 24-bit significand, shifted right and
// s = 0 for values below 2^-38; bits below 2^-62 are dropped).  Adding words
// is integer addition, exact and order-free, so every vertex's sum is the
// exact sum of its FP32 contributions (each truncated below 2^-62) whatever
// the schedule, and runs are bit-reproducible.  A row sum is at most the
// total rank (1), far below the 4 that 64 bits with 62 fraction bits hold.
#include <algorithm>
#include <numeric>

#include "graph_internal.hpp"
#include "pr_common.hpp"

#include <rocprim/rocprim.hpp>

namespace ghet {
namespace {

constexpr int kWinShift = 15;
constexpr int kWin = 1 << kWinShift;      // sources per window (128 KB of words in LDS)
constexpr int kBinMax = 1 << 14;          // destinations per bin (128 KB of u64 in LDS)
constexpr int kMaxWin = 2048;             // windows (n <= 2^26)
constexpr int kMaxSeg = 3072;             // bins a gather item's LDS segment table holds
constexpr int64_t kBinEdges = 1 << 17;    // in-edges a bin aims for
constexpr int64_t kItemEdges = 1 << 17;   // edges per gather item
constexpr int kPB = 1024;                 // threads per block, both kernels (one block per CU)
constexpr int kWaves = kPB / 64;
constexpr int kQ = 8;                     // rows in flight per thread (gather)

// FP32 x in [0, 2) -> contribution word; x = m * 2^(ex - 150) for a normal x.
__device__ __forceinline__ uint32_t fix_pack(float x) {
    const uint32_t bits = __float_as_uint(x);
    const int ex = static_cast<int>(bits >> 23);
    if (ex == 0) return 0u;  // zero or subnormal: below 2^-126, under the resolution
    uint32_t mant = (bits & 0x7FFFFFu) | 0x800000u;
    int sh = ex - 88;        // x * 2^62 = mant * 2^(ex - 88)
    if (sh < 0) {
        mant = sh > -24 ? mant >> -sh : 0u;
        sh = 0;
    }
    return (static_cast<uint32_t>(sh) << 24) | mant;
}

// Contribution word -> x * 2^62 as an integer (below 2^63 for x < 2).
__device__ __forceinline__ unsigned long long fix_unpack(uint32_t w) {
    return static_cast<unsigned long long>(w & 0xFFFFFFu) << (w >> 24);
}

__global__ void __launch_bounds__(256) k_prb_init(const int32_t* __restrict__ outdeg, int32_t n, double pr0,
                                                  uint32_t* __restrict__ contrib, double* __restrict__ part) {
    __shared__ double s_red[256 / 64];
    double sink = 0;
    for (int64_t v = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; v < n;
         v += static_cast<int64_t>(gridDim.x) * 256) {
        const int32_t d = outdeg[v];
        if (d > 0) contrib[v] = fix_pack(static_cast<float>(pr0 / d));
        else {
            contrib[v] = 0u;
            sink += pr0;
        }
    }
    const double t = block_sum_f64<256>(sink, s_red);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// Last c in [lo, hi) with s[c] <= k, given s[lo] <= k < s[hi]; k is the same
// in every lane (the reads are broadcasts and the loop is uniform).
__device__ __forceinline__ int prb_search(const int32_t* s, int32_t k, int lo, int hi) {
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (__builtin_amdgcn_readfirstlane(s[mid]) <= k) lo = mid;
        else hi = mid;
    }
    return lo;
}

// One block per item {w, e0, e1, b0}.  Each wave takes a contiguous run of
// 64-edge rows of [e0, e1) and walks it with its current segment c (bin
// b0 + c): a row inside the segment stores at e + D(c); a row that reaches
// the next segment start steps c forward (a uniform loop; lanes at or past
// each start take that segment's D).  kQ rows' positions are formed first,
// then their source loads are issued together.
__global__ void __launch_bounds__(kPB) k_prb_gather(const int4* __restrict__ items, const int32_t* __restrict__ item_nseg,
                                                    const int2* __restrict__ segtab, int32_t num_bin,
                                                    const uint16_t* __restrict__ src, const uint32_t* __restrict__ contrib,
                                                    int32_t n, int32_t spare, uint32_t* __restrict__ val) {
    __shared__ uint32_t s_win[kWin];
    __shared__ int32_t s_S[kMaxSeg + 1];  // window-major start of segment b0 + c (s_S[nseg]: past the item)
    __shared__ int32_t s_D[kMaxSeg];      // bin-major start minus window-major start
    const int4 it = items[blockIdx.x];
    const int32_t nseg = item_nseg[blockIdx.x];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int64_t w0 = static_cast<int64_t>(it.x) << kWinShift;
    const int32_t wn = static_cast<int32_t>(n - w0 < kWin ? n - w0 : kWin);
    {
        // contrib + w0 is 16-byte aligned: w0 is a multiple of 2^15
        const int32_t wn4 = wn >> 2;
        const uint4* __restrict__ c4 = reinterpret_cast<const uint4*>(contrib + w0);
        uint4* s4 = reinterpret_cast<uint4*>(s_win);
        for (int32_t i = tid; i < wn4; i += kPB) s4[i] = c4[i];
        for (int32_t i = (wn4 << 2) + tid; i < wn; i += kPB) s_win[i] = contrib[w0 + i];
        const int2* st = segtab + static_cast<int64_t>(it.x) * num_bin + it.w;
        for (int32_t i = tid; i <= nseg; i += kPB) {
            const int2 d = st[i];
            s_S[i] = d.x;
            if (i < nseg) s_D[i] = d.y - d.x;
        }
    }
    __syncthreads();
    const int32_t e0 = it.y, e1 = it.z;
    const int32_t R = (e1 - e0 + 63) >> 6;
    const int32_t r0 = static_cast<int32_t>(static_cast<int64_t>(R) * wave / kWaves);
    const int32_t r1 = static_cast<int32_t>(static_cast<int64_t>(R) * (wave + 1) / kWaves);
    const int32_t a = __builtin_amdgcn_readfirstlane(e0 + 64 * r0);
    const int32_t z = __builtin_amdgcn_readfirstlane(min(e1, e0 + 64 * r1));
    if (a >= z) return;
    int c = prb_search(s_S, a, 0, nseg);  // s_S[0] <= e0 <= a < e1 <= s_S[nseg]
    int32_t D = __builtin_amdgcn_readfirstlane(s_D[c]);
    int32_t nb = __builtin_amdgcn_readfirstlane(s_S[c + 1]);
    for (int32_t k0 = a; k0 < z; k0 += 64 * kQ) {
        int32_t pos[kQ], ee[kQ];
#pragma unroll
        for (int j = 0; j < kQ; ++j) {
            const int32_t r = k0 + 64 * j;
            const int32_t e = r + lane;
            int32_t p = e + D;
            if (r < z) {
                const int32_t rl = min(r + 63, z - 1);
                while (nb <= rl) {  // the row reaches segment c + 1 (c + 1 < nseg: nb <= rl < e1 <= s_S[nseg])
                    ++c;
                    const int32_t Dn = __builtin_amdgcn_readfirstlane(s_D[c]);
                    p = e >= nb ? e + Dn : p;
                    D = Dn;
                    nb = __builtin_amdgcn_readfirstlane(s_S[c + 1]);
                }
            }
            // lanes past the run load edge e0 and store to the spare slot
            // val[m], so the loads and stores need no branches
            pos[j] = e < z ? p : spare;
            ee[j] = e < z ? e : e0;
        }
        uint32_t si[kQ];
#pragma unroll
        for (int j = 0; j < kQ; ++j) si[j] = src[ee[j]];
#pragma unroll
        for (int j = 0; j < kQ; ++j) val[pos[j]] = s_win[si[j]];
    }
}

// One block per bin (bins in descending edge order): the bin's edges are
// [binoff[b], binoff[b+1]) in bin-major order.  Each thread keeps two steps
// of kQG 4-element groups in flight (one 8-byte and one 16-byte load per
// group, against four 2-byte and four 4-byte loads: 442 -> 425 us at s24):
// the next step's loads are issued before the current step's values are
// added.
__global__ void __launch_bounds__(kPB) k_prb_accum(
    const int32_t* __restrict__ bin_order, const int32_t* __restrict__ bin_start, const int32_t* __restrict__ binoff,
    const uint16_t* __restrict__ dst, const uint32_t* __restrict__ val, const int32_t* __restrict__ outdeg,
    const PRScalars* __restrict__ scal, double nd, double damping, uint32_t* __restrict__ next,
    double* __restrict__ part, double* __restrict__ rank_out, int last) {
    __shared__ unsigned long long s_acc[kBinMax];
    __shared__ double s_red[kWaves];
    const int tid = threadIdx.x;
    const int32_t b = bin_order[blockIdx.x];
    const int32_t v0 = bin_start[b], nv = bin_start[b + 1] - v0;
    const int32_t base = binoff[b], T = binoff[b + 1] - base;
    for (int32_t i = tid; i < nv; i += kPB) s_acc[i] = 0ull;
    __syncthreads();
    // groups of 4 elements aligned at multiples of 4: one 8-byte load of
    // destinations and one 16-byte load of values per group; elements
    // outside [base, base + T) are skipped (the span's two end groups read
    // at most 3 elements either side, inside the layout's arena)
    constexpr int kQG = 2;
    const int32_t g0 = base >> 2, G = ((base + T + 3) >> 2) - g0;
    const uint2* __restrict__ d4 = reinterpret_cast<const uint2*>(dst);
    const uint4* __restrict__ v4 = reinterpret_cast<const uint4*>(val);
    constexpr int32_t kStepG = kPB * kQG;
    uint2 dd[kQG];
    uint4 vv[kQG];
#pragma unroll
    for (int j = 0; j < kQG; ++j) {
        const int32_t g = j * kPB + tid;
        const int32_t gi = g < G ? g0 + g : g0;
        dd[j] = d4[gi];
        vv[j] = v4[gi];
    }
    for (int32_t k0 = 0; k0 < G; k0 += kStepG) {
        uint2 dn[kQG];
        uint4 vn[kQG];
#pragma unroll
        for (int j = 0; j < kQG; ++j) {
            const int32_t g = k0 + kStepG + j * kPB + tid;
            const int32_t gi = g < G ? g0 + g : g0;
            dn[j] = d4[gi];
            vn[j] = v4[gi];
        }
#pragma unroll
        for (int j = 0; j < kQG; ++j) {
            const int32_t g = k0 + j * kPB + tid;
            if (g < G) {
                const int32_t e = (g0 + g) << 2;
                const uint32_t dq[4] = {dd[j].x & 0xFFFFu, dd[j].x >> 16, dd[j].y & 0xFFFFu, dd[j].y >> 16};
                const uint32_t vq[4] = {vv[j].x, vv[j].y, vv[j].z, vv[j].w};
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (e + q >= base && e + q < base + T) atomicAdd(&s_acc[dq[q] & (kBinMax - 1)], fix_unpack(vq[q]));
            }
        }
#pragma unroll
        for (int j = 0; j < kQG; ++j) {
            dd[j] = dn[j];
            vv[j] = vn[j];
        }
    }
    __syncthreads();
    const double base_pr = scal->sink * damping / nd + (1.0 - damping) / nd;
    double sink = 0;
    for (int32_t i = tid; i < nv; i += kPB) {
        const double pr = base_pr + damping * (static_cast<double>(s_acc[i]) * 0x1p-62);
        const int32_t d = outdeg[v0 + i];
        if (d > 0) next[v0 + i] = fix_pack(static_cast<float>(pr / d));
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

// Segment lengths in bin-major order: len[b * num_win + w] = |segment (w, b)|.
__global__ void k_prb_lens(const int32_t* __restrict__ soff, int32_t num_win, int32_t num_bin, int32_t* __restrict__ len) {
    const int64_t total = static_cast<int64_t>(num_win) * num_bin;
    for (int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; q < total;
         q += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t b = q / num_win, w = q - b * num_win;
        const int64_t k = w * num_bin + b;
        len[q] = soff[k + 1] - soff[k];
    }
}

// segtab[w * num_bin + b] = {window-major start, bin-major start}, segtab[K] =
// {m, m}; binoff[b] = bin-major start of bin b (binoff[num_bin] = m);
// win[w] = first edge of window w (win[num_win] = m).
__global__ void k_prb_segtab(const int32_t* __restrict__ soff, const int32_t* __restrict__ pos_bm, int32_t num_win,
                             int32_t num_bin, int64_t m, int2* __restrict__ segtab, int32_t* __restrict__ binoff,
                             int32_t* __restrict__ win) {
    const int64_t total = static_cast<int64_t>(num_win) * num_bin;
    for (int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; q <= total;
         q += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t w = q / num_bin, b = q - w * num_bin;
        if (q < total) segtab[q] = make_int2(soff[q], pos_bm[b * num_win + w]);
        else segtab[q] = make_int2(static_cast<int32_t>(m), static_cast<int32_t>(m));
        if (b == 0) win[w] = soff[q];
        if (q < num_bin) binoff[q] = pos_bm[q * num_win];
        if (q == num_bin) binoff[q] = static_cast<int32_t>(m);
    }
}

__global__ void k_prb_src(const uint32_t* __restrict__ word, int64_t m, uint16_t* __restrict__ src) {
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < m;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x)
        src[e] = static_cast<uint16_t>(word[e] & 0xFFFFu);
}

// Destinations to bin-major order: one wave per segment.
__global__ void k_prb_dst(const uint32_t* __restrict__ word, const int2* __restrict__ segtab, int64_t num_seg,
                          uint16_t* __restrict__ dst) {
    const int lane = lane_id();
    const int64_t wave = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const int64_t waves = (static_cast<int64_t>(gridDim.x) * blockDim.x) >> 6;
    for (int64_t q = wave; q < num_seg; q += waves) {
        const int2 a = segtab[q];
        const int32_t len = segtab[q + 1].x - a.x;  // window-major segments are contiguous
        for (int32_t i = lane; i < len; i += 64) dst[a.y + i] = static_cast<uint16_t>(word[a.x + i] >> 16);
    }
}

// Per gather item: the bin of its first edge's segment (it.w) and the number
// of segments its edges touch.
__global__ void k_prb_item_segs(int4* __restrict__ items, int32_t num_items, const int2* __restrict__ segtab,
                                int32_t num_bin, int32_t* __restrict__ nseg) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < num_items;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        int4 it = items[i];
        const int2* st = segtab + static_cast<int64_t>(it.x) * num_bin;
        auto last_le = [&](int32_t e) {  // last b in [0, num_bin) with st[b].x <= e
            int lo = 0, hi = num_bin;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (st[mid].x <= e) lo = mid;
                else hi = mid;
            }
            return lo;
        };
        const int b0 = last_le(it.y), b1 = last_le(it.z - 1);
        it.w = b0;
        items[i] = it;
        nseg[i] = b1 - b0 + 1;
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

// Bins: consecutive destinations, closed before a bin would pass kBinMax
// vertices or kBinEdges in-edges (a vertex with more in-edges than that is a
// bin of its own).
std::vector<int32_t> prb_bins(const std::vector<int32_t>& rp, int32_t n, std::vector<int64_t>* edges) {
    std::vector<int32_t> bs{0};
    int64_t e = 0;
    int32_t start = 0;
    for (int32_t v = 0; v < n; ++v) {
        const int64_t d = rp[v + 1] - rp[v];
        if (v > start && (v - start == kBinMax || e + d > kBinEdges)) {
            bs.push_back(v);
            edges->push_back(e);
            start = v;
            e = 0;
        }
        e += d;
    }
    bs.push_back(n);
    edges->push_back(e);
    return bs;
}

int build_prb_layout(ghet_graph* g, hipStream_t s) {
    PRBinLayout& L = g->prb;
    if (L.built) return GHET_OK;
    const int32_t n = g->n;
    const int64_t m = g->in.m;
    if (!g->out.present() || g->out.col == nullptr || g->out.m != m)
        return set_error(GHET_ERR_INVALID, "PageRank bins need the out-CSR with the in-CSR's edges");
    if (!pr_binned_fits(n, m)) return set_error(GHET_ERR_UNSUPPORTED, "PageRank bins: graph too large for the bin tables");
    std::vector<int64_t> bedges;
    const std::vector<int32_t> bs = prb_bins(g->h_in_rowptr, n, &bedges);
    const int32_t NB = static_cast<int32_t>(bs.size() - 1);
    const int32_t NW = static_cast<int32_t>(ceil_div(n, kWin));
    const int64_t K = int64_t(NW) * NB;
    if (NB > kMaxSeg || K >= INT32_MAX)
        return set_error(GHET_ERR_UNSUPPORTED, "PageRank bins: %d bins over %d windows", NB, NW);
    std::vector<int32_t> order(NB);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return bedges[a] > bedges[b]; });
    const int64_t max_items = m / kItemEdges + NW + 1;

    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t me = std::max<size_t>(size_t(m), 1);
    const size_t bytes = 2 * al(2 * me) + al(4 * (me + 1)) + al(8 * size_t(K + 1)) + 2 * al(4 * (size_t(NB) + 1)) +
                         al(4 * size_t(NB)) + al(16 * size_t(max_items)) + al(4 * size_t(max_items)) + al(4 * size_t(n));
    GHET_HIP(hipMalloc(&L.mem, bytes));
    {
        char* p = static_cast<char*>(L.mem);
        auto take = [&](size_t b) { char* r = p; p += al(b); return r; };
        L.src = reinterpret_cast<uint16_t*>(take(2 * me));
        L.dst = reinterpret_cast<uint16_t*>(take(2 * me));
        L.val = reinterpret_cast<uint32_t*>(take(4 * (me + 1)));  // val[m]: the gather's spare slot
        L.segtab = reinterpret_cast<int2*>(take(8 * size_t(K + 1)));
        L.binoff = reinterpret_cast<int32_t*>(take(4 * (size_t(NB) + 1)));
        L.bin_start = reinterpret_cast<int32_t*>(take(4 * (size_t(NB) + 1)));
        L.bin_order = reinterpret_cast<int32_t*>(take(4 * size_t(NB)));
        L.items = reinterpret_cast<int4*>(take(16 * size_t(max_items)));
        L.item_nseg = reinterpret_cast<int32_t*>(take(4 * size_t(max_items)));
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
        int32_t* len = t.get<int32_t>(size_t(K) + 1);
        int32_t* pos = t.get<int32_t>(size_t(K) + 1);
        int32_t* aux = t.get<int32_t>(size_t(scan_aux_ints(K)));
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
        hipLaunchKernelGGL(k_prb_lens, dim3(grid_for(K)), dim3(256), 0, s, soff, NW, NB, len);
        ok(hipGetLastError());
        if (t.err == hipSuccess && exclusive_scan_i32(len, pos, K, aux, s) != GHET_OK) t.err = hipErrorUnknown;
        hipLaunchKernelGGL(k_prb_segtab, dim3(grid_for(K + 1)), dim3(256), 0, s, soff, pos, NW, NB, m, L.segtab, L.binoff,
                           d_win);
        hipLaunchKernelGGL(k_prb_src, dim3(grid_for(m)), dim3(256), 0, s, wd.current(), m, L.src);
        hipLaunchKernelGGL(k_prb_dst, dim3(grid_for(K * 64)), dim3(256), 0, s, wd.current(), L.segtab, K, L.dst);
        ok(hipGetLastError());
        ok(hipMemcpyAsync(win.data(), d_win, 4 * win.size(), hipMemcpyDeviceToHost, s));
        ok(hipStreamSynchronize(s));
        if (t.err != hipSuccess) return fail("build", t.err);
    }
    // Gather items: each window's edges cut into ranges of at most
    // kItemEdges; the window offsets are checked here once.
    std::vector<int4> items;
    bool good = win[0] == 0 && win[NW] == m;
    for (int32_t w = 0; good && w < NW; ++w) {
        good = win[w] <= win[w + 1];
        for (int64_t e0 = win[w]; good && e0 < win[w + 1]; e0 += kItemEdges)
            items.push_back(make_int4(w, static_cast<int32_t>(e0),
                                      static_cast<int32_t>(std::min<int64_t>(win[w + 1], e0 + kItemEdges)), 0));
    }
    if (!good || static_cast<int64_t>(items.size()) > max_items) {
        hipFree(L.mem);
        L = PRBinLayout{};
        return set_error(GHET_ERR_INVALID, "PageRank bins: window offsets failed their check");
    }
    std::vector<int32_t> nseg(items.size());
    if (!items.empty()) {
        hipError_t e = hipMemcpyAsync(L.items, items.data(), 16 * items.size(), hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return fail("items", e);
        hipLaunchKernelGGL(k_prb_item_segs, dim3(grid_for(int64_t(items.size()))), dim3(256), 0, s, L.items,
                           static_cast<int32_t>(items.size()), L.segtab, NB, L.item_nseg);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(items.data(), L.items, 16 * items.size(), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipMemcpyAsync(nseg.data(), L.item_nseg, 4 * nseg.size(), hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) return fail("items", e);
    }
    const hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail("items", e);
    // each item's segments fit the gather kernel's LDS table and stay inside
    // the window's part of the segment table
    for (size_t i = 0; i < items.size() && good; ++i)
        good = items[i].w >= 0 && nseg[i] >= 1 && items[i].w + nseg[i] <= NB && nseg[i] <= kMaxSeg;
    if (!good) {
        hipFree(L.mem);
        L = PRBinLayout{};
        return set_error(GHET_ERR_INVALID, "PageRank bins: item segments failed their check");
    }
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

// From 2^19 vertices on the bins win.  Measured per iteration (ms, binned
// against gathered), R-MAT ef 16: s19 0.059 / 0.069, s20 0.068 / 0.092,
// s21 0.129 / 0.177, s22 0.259 / 0.377, s24 0.889 / 1.62; Erdos-Renyi with
// the s24 sizes 0.861 / 4.91 (no hot sources for the gathers to keep).
bool pr_binned_default(int32_t n, int64_t m) {
    return n >= (1 << 19) && pr_binned_fits(n, m);
}

int run_pr_binned(ghet_graph* g, int32_t iters, double damping, double* pr_out, bool normalise, const ghet_opts* o,
                  ghet_run_stats* st) {
    const int32_t n = g->n;
    hipStream_t s = pick_stream(g, o);
    GHET_CHECK(build_prb_layout(g, s));
    const PRBinLayout& L = g->prb;
    const int32_t init_blocks = static_cast<int32_t>(std::min<int64_t>(ceil_div(n, 256), 4096));
    void* p;
    uint32_t *ca, *cb;
    double *rank, *out, *part;
    PRScalars* scal;
    GHET_CHECK(g->ws.get(WS_PR_CONTRIB, 4 * size_t(n), &p)); ca = static_cast<uint32_t*>(p);
    GHET_CHECK(g->ws.get(WS_LABEL_A, 4 * size_t(n), &p)); cb = static_cast<uint32_t*>(p);
    GHET_CHECK(g->ws.get(WS_PR_RANK, 8 * size_t(n), &p)); rank = static_cast<double*>(p);
    GHET_CHECK(g->ws.get(WS_PR_OUT, 8 * size_t(n), &p)); out = static_cast<double*>(p);
    GHET_CHECK(g->ws.get(WS_PR_PART, 8 * size_t(std::max(L.num_bin, init_blocks) + 1), &p)); part = static_cast<double*>(p);
    GHET_CHECK(g->ws.get(WS_COUNTERS, 256, &p)); scal = static_cast<PRScalars*>(p);
    const double inv_n = 1.0 / n;
    hipEvent_t* kev = nullptr;
    if (st) GHET_CHECK(kernel_events(g, 3 * iters, &kev));

    GHET_HIP(hipEventRecord(g->ev_begin, s));
    hipLaunchKernelGGL(k_prb_init, dim3(init_blocks), dim3(256), 0, s, L.outdeg, n, inv_n, ca, part);
    hipLaunchKernelGGL(k_pr_reduce, dim3(1), dim3(1024), 0, s, part, init_blocks, &scal->sink, static_cast<int*>(nullptr));
    if (iters == 0) GHET_CHECK(fill<double>(rank, inv_n, n, s));
    for (int32_t it = 0; it < iters; ++it) {
        const int last = (it == iters - 1);
        if (st) GHET_HIP(hipEventRecord(kev[3 * it], s));
        if (L.num_items > 0)
            hipLaunchKernelGGL(k_prb_gather, dim3(L.num_items), dim3(kPB), 0, s, L.items, L.item_nseg, L.segtab, L.num_bin,
                               L.src, ca, n, static_cast<int32_t>(g->in.m), L.val);
        if (st) GHET_HIP(hipEventRecord(kev[3 * it + 1], s));
        hipLaunchKernelGGL(k_prb_accum, dim3(L.num_bin), dim3(kPB), 0, s, L.bin_order, L.bin_start, L.binoff, L.dst, L.val,
                           L.outdeg, scal, static_cast<double>(n), damping, cb, part, rank, last);
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
