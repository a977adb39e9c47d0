// PageRank on gfx950: pull gather over the in-CSR in FP32 with compensated
// per-vertex sums; dangling mass, the stored rank and the final
// normalisation in FP64; every reduction in a fixed order (deterministic).
//
// Reference: GPU_PR (GPU_PR_csr.cuh:34-117) runs one thread per vertex over
// its in-list in FP64 (hub vertices serialise) and sums dangling rank with a
// CAS-loop atomicAdd (order nondeterministic).  Semantics of PageRank
// (CPU_PageRank.hpp:12-56): pr0 = 1/N; each iteration
//     pr'[v] = (S*d/N + (1-d)/N) + d * sum_{u in in(v)} pr[u]/outdeg(u),
// S = sum of pr over vertices without out-edges.
//
// Relabeling: the layout renumbers vertices by out-degree, descending (a
// counting sort, stable by id), and stores the in-CSR in that numbering.  The
// gathered contribution array is then hot at its front: on R-MAT the top 2^20
// sources (4 MB of FP32, one XCD's L2) are the targets of ~98% of gathers at
// scale 22, against ~57% in the original numbering.  Results are written back
// in the caller's numbering by a final scatter.
//
// Layout (built once per graph from the in-offsets): each row's sources are
// split into a hot part (relabeled ids below `hot`) and cold segments (one per
// source window).  The row space is the n hot parts followed by the segments
// (window-major); consecutive short entries are packed into tiles of at most
// kTileNnz edges and kTileRows entries, and an entry with more than kTileNnz
// edges is split into kTileNnz-edge chunks.  Chunks and tiles form one list
// of 32-byte work items: the hot phase first, then one group per window.
// One launch per iteration walks the list with persistent blocks:
//   1. gather contrib[col[e]] for the item's contiguous edge range (8 per
//      lane, all issued before the first wait; the next item's column, row
//      offset, out-degree and slot loads are issued behind them);
//   2. tiles stage the values in LDS -- rows with <= 32 values: one lane
//      each; longer rows: one wave; chunks reduce in the block;
//   3. a row without partials writes contrib'[v] = (float)(pr'[v] /
//      outdeg(v)) for the next iteration, folds pr'[v] of sinks into a
//      per-thread FP64 sum, and on the last iteration writes pr'[v] in FP64;
//      chunks, cold segments and the hot parts of rows with cold sources
//      write FP64 partials to their slots instead.
// k_pr_long then finishes the rows with partials (slots summed in order), and
// a one-block kernel sums the sink partials in index order into S.
#include <algorithm>
#include <atomic>
#include <cmath>

#include "graph_internal.hpp"
#include "pr_common.hpp"

#include <rocprim/rocprim.hpp>

namespace ghet {
namespace {

constexpr int kBT = 256;
constexpr int kTileNnz = 2048;
constexpr int kTileRows = 512;
constexpr int kStage = kTileNnz / kBT;    // staged values per thread
constexpr int kRowsPT = kTileRows / kBT;  // tile rows per thread
constexpr int kWaveRow = 32;            // rows with more staged values than this use a wave
// Source-blocked hot phase (k_pr_hot)
constexpr int kHotShift = 14;
constexpr int kHotB = 1 << kHotShift;   // hot sources per LDS block (64 KB of FP32)
constexpr int kHotT = 8192;             // rows per tile (two FP32 accumulators each: 64 KB)
constexpr int kHotBT = 1024;            // threads per block (one block per CU)
constexpr int kHotK = 8;                // consecutive slice words per thread and chunk
constexpr int kHotFillMin = 2048;       // slices with fewer words gather from global memory
constexpr int kHotChunk = kHotBT * kHotK;
// Row sums are accumulated as unsigned fixed point with 60 fraction bits: a
// row's sum of contributions is at most the total rank (1), integer adds are
// exact and order-free (so the sums are deterministic), and the resolution
// 2^-60 is far below the FP32 rounding of the contributions themselves.
constexpr double kHotFix = 1152921504606846976.0;  // 2^60
static_assert(kHotT <= (1 << (32 - kHotShift)), "row in tile must fit beside the source in a word");

// Compensated wave reduction of per-lane Kahan accumulators (fixed tree order).
__device__ __forceinline__ Kahan32 wave_kahan(Kahan32 k) {
    for (int off = 32; off > 0; off >>= 1) {
        Kahan32 o;
        o.s = __shfl_xor(k.s, off);
        o.c = __shfl_xor(k.c, off);
        // combine in a lane-independent order so every lane holds the same value
        Kahan32 lo = (lane_id() & off) ? o : k, hi = (lane_id() & off) ? k : o;
        lo.merge(hi);
        k = lo;
    }
    return k;
}

// A work item: a tile of whole rows (row_end > row) or one kTileNnz-edge
// chunk of a long row (row_end < 0).  32 bytes, read with scalar loads.
struct alignas(32) PRItem {
    int32_t row;      // first row of the tile / the long row
    int32_t row_end;  // one past the tile's last row; -1 for a chunk
    int32_t e0, e1;   // edge range [e0, e1) in the relabeled in-CSR
    int32_t chunk;    // chunk: global chunk index (slot in long_part)
    int32_t pad[3];
};

// Register-resident loads of one work item: the staged column indices (lanes
// past the end hold 0, a valid index, and are masked later), and for tiles
// the row offsets, out-degrees and partial slots of rows tid + j*kBT.
struct PRStage {
    int32_t idx[kStage];
    int32_t roff[kRowsPT];
    int32_t od[kRowsPT];
    int32_t sl[kRowsPT];
};

__device__ __forceinline__ void pr_stage(const PRItem& it, bool valid, const int32_t* __restrict__ rowptr,
                                         const int32_t* __restrict__ col, const int32_t* __restrict__ outdeg,
                                         const int32_t* __restrict__ slot, int32_t n, PRStage& st) {
    const int tid = threadIdx.x;
    const int32_t nnz = valid ? it.e1 - it.e0 : 0;
    // Branch-free: lanes past the end load element 0 (always allocated) and
    // are masked where the values are used, so every load is unconditional
    // and the compiler can wait for the gathers alone.
#pragma unroll
    for (int j = 0; j < kStage; ++j) {
        const int k = tid + j * kBT;
        st.idx[j] = col[k < nnz ? it.e0 + k : 0];
    }
    const int32_t rows = (valid && it.row_end >= 0) ? it.row_end - it.row : 0;
#pragma unroll
    for (int j = 0; j < kRowsPT; ++j) {
        const int r = tid + j * kBT;
        const int32_t v = r < rows ? it.row + r : 0;
        st.roff[j] = rowptr[v];
        st.od[j] = outdeg[v < n ? v : 0];  // segments (v >= n) never finish a vertex
        st.sl[j] = slot[v];
    }
}

// One iteration over the relabeled in-CSR.  Persistent blocks walk the item
// list with stride gridDim.x; the loads of item i+G (columns, row offsets,
// out-degrees, slots) are issued while the gathers of item i are in flight,
// so a block always has two items' worth of memory requests outstanding and
// no item pays more than one exposed round trip.  A tile row with slot -1
// finishes its vertex; any other tile row (the hot part of a vertex with cold
// sources, or a cold segment) and every chunk store a partial sum for
// k_pr_long.  Every thread keeps its sink mass in item order;
// part[blockIdx.x] gets the block's total.  Launch bounds ask for 8 waves per
// SIMD (64 VGPRs, no spills): 1691 -> 1632 us per iteration at s24 against
// the 5 waves the default allocation (82 VGPRs) allowed.
__global__ void __launch_bounds__(kBT, 8) k_pr_iter(
    const PRItem* __restrict__ items, int32_t num_items, const int32_t* __restrict__ rowptr,
    const int32_t* __restrict__ col, const int32_t* __restrict__ outdeg, const int32_t* __restrict__ slot, int32_t n,
    double* __restrict__ long_part, const float* __restrict__ contrib, float* __restrict__ next,
    const PRScalars* __restrict__ scal, double nd, double damping, double* __restrict__ part,
    double* __restrict__ rank_out, int last) {
    __shared__ float s_val[kTileNnz];
    __shared__ int32_t s_off[kTileRows + 1];
    __shared__ int32_t s_od[kTileRows];
    __shared__ int32_t s_sl[kTileRows];
    __shared__ double s_red[kBT / 64];
    __shared__ float s_ks[kBT / 64], s_kc[kBT / 64];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const double base = scal->sink * damping / nd + (1.0 - damping) / nd;
    const int32_t G = gridDim.x;
    double sink = 0;
    int32_t i = blockIdx.x;
    PRItem cur = i < num_items ? items[i] : PRItem{};
    PRStage st;
    pr_stage(cur, i < num_items, rowptr, col, outdeg, slot, n, st);
    PRItem nxt = i + G < num_items ? items[i + G] : PRItem{};
    for (; i < num_items; i += G) {
        float val[kStage];
#pragma unroll
        for (int j = 0; j < kStage; ++j) val[j] = contrib[st.idx[j]];
        // the next item's loads overlap these gathers
        PRStage sn;
        pr_stage(nxt, i + G < num_items, rowptr, col, outdeg, slot, n, sn);
        const PRItem nxt2 = i + 2 * G < num_items ? items[i + 2 * G] : PRItem{};
        const int32_t nnz = cur.e1 - cur.e0;
        __syncthreads();  // the previous item's LDS readers are done
        if (cur.row_end >= 0) {
            const int32_t rows = cur.row_end - cur.row;
#pragma unroll
            for (int j = 0; j < kStage; ++j) {
                const int k = tid + j * kBT;
                if (k < nnz) s_val[k] = val[j];
            }
#pragma unroll
            for (int j = 0; j < kRowsPT; ++j) {
                const int r = tid + j * kBT;
                if (r < rows) {
                    s_off[r] = st.roff[j] - cur.e0;
                    s_od[r] = st.od[j];
                    s_sl[r] = st.sl[j];
                }
            }
            if (tid == 0) s_off[rows] = nnz;
            __syncthreads();
            // rows with few values: one lane each, in edge order
#pragma unroll
            for (int j = 0; j < kRowsPT; ++j) {
                const int r = tid + j * kBT;
                if (r < rows) {
                    const int32_t a = s_off[r], z = s_off[r + 1];
                    if (z - a <= kWaveRow) {
                        Kahan32 acc;
                        for (int32_t e = a; e < z; ++e) acc.add(s_val[e]);
                        if (st.sl[j] >= 0) {
                            long_part[st.sl[j]] = static_cast<double>(acc.s) + static_cast<double>(acc.c);
                        } else {
                            const double pr = base + damping * static_cast<double>(acc.value());
                            if (st.od[j] > 0) next[cur.row + r] = static_cast<float>(pr / st.od[j]);
                            else sink += pr;
                            if (last) rank_out[cur.row + r] = pr;
                        }
                    }
                }
            }
            // rows with more values: wave w takes 64-row groups w, w+4, ...
            for (int c0 = wave * 64; c0 < rows; c0 += kBT) {
                const int r = c0 + lane;
                const bool wide = r < rows && (s_off[r + 1] - s_off[r]) > kWaveRow;
                uint64_t mask = __ballot(wide);
                while (mask) {
                    const int rr = c0 + __ffsll(static_cast<unsigned long long>(mask)) - 1;
                    mask &= mask - 1;
                    const int32_t a = s_off[rr], z = s_off[rr + 1];
                    Kahan32 acc;
                    for (int32_t e = a + lane; e < z; e += 64) acc.add(s_val[e]);
                    acc = wave_kahan(acc);
                    if (lane == 0) {
                        const int32_t sl = s_sl[rr];
                        if (sl >= 0) {
                            long_part[sl] = static_cast<double>(acc.s) + static_cast<double>(acc.c);
                        } else {
                            const double pr = base + damping * static_cast<double>(acc.value());
                            const int32_t d = s_od[rr];
                            if (d > 0) next[cur.row + rr] = static_cast<float>(pr / d);
                            else sink += pr;
                            if (last) rank_out[cur.row + rr] = pr;
                        }
                    }
                }
            }
        } else {
            // chunk of a long row: compensated block sum, then the last chunk
            // to arrive combines the row's partials in chunk order
            Kahan32 acc;
#pragma unroll
            for (int j = 0; j < kStage; ++j)
                if (tid + j * kBT < nnz) acc.add(val[j]);
            acc = wave_kahan(acc);
            if (lane == 0) {
                s_ks[wave] = acc.s;
                s_kc[wave] = acc.c;
            }
            __syncthreads();
            if (tid == 0) {
                Kahan32 t;
                for (int w = 0; w < kBT / 64; ++w) {
                    Kahan32 o;
                    o.s = s_ks[w];
                    o.c = s_kc[w];
                    t.merge(o);
                }
                long_part[cur.chunk] = static_cast<double>(t.s) + static_cast<double>(t.c);
            }
        }
        cur = nxt;
        nxt = nxt2;
        st = sn;
    }
    const double t = block_sum_f64(sink, s_red);
    if (tid == 0) part[blockIdx.x] = t;
}

// Rows finished from partials (long rows and rows with cold sources): sum
// each row's slots in slot order and finish the vertex (one thread per row).
// Runs after k_pr_iter in stream order, which makes every partial visible
// without device-scope fences -- a fence per partial would write back and
// invalidate the XCD's L2 and with it the hot contributions.  Sink mass goes
// to part[blockIdx.x].
__global__ void __launch_bounds__(kBT) k_pr_long(const int4* __restrict__ long_rows, int32_t num_long_rows,
                                                 const double* __restrict__ long_part,
                                                 const int32_t* __restrict__ outdeg, float* __restrict__ next,
                                                 const PRScalars* __restrict__ scal, double nd, double damping,
                                                 double* __restrict__ part, double* __restrict__ rank_out, int last) {
    __shared__ double s_red[kBT / 64];
    const double base = scal->sink * damping / nd + (1.0 - damping) / nd;
    double sink = 0;
    for (int32_t L = blockIdx.x * kBT + threadIdx.x; L < num_long_rows; L += gridDim.x * kBT) {
        const int4 lr = long_rows[L];  // {row, first slot, slots, 0}
        double total = 0;
        for (int32_t j = 0; j < lr.z; ++j) total += long_part[lr.y + j];
        const double pr = base + damping * total;
        const int32_t d = outdeg[lr.x];
        if (d > 0) next[lr.x] = static_cast<float>(pr / d);
        else sink += pr;
        if (last) rank_out[lr.x] = pr;
    }
    const double t = block_sum_f64(sink, s_red);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// Source-blocked hot phase: one block (1024 threads, one per CU) per row
// tile (at most kHotT rows) at a time, tiles handed out by an atomic counter
// in row order (the heavy tiles first).  For each source block b the block
// copies contrib[b kHotB, (b+1) kHotB) into LDS with coalesced loads, then
// walks the slice (tile, b), whose words are sorted by row: every gather of
// a hot source is an LDS read instead of a random global load (the per-edge
// gather phase is bound by L2 requests).  Each thread takes kHotK
// consecutive words, sums each run of equal rows in FP64 and adds it to the
// row's LDS accumulator as a 64-bit fixed-point integer (ds_add_u64): integer
// addition is exact and commutative, so the result does not depend on which
// thread adds first.  After the last block the tile's rows are finished as
// k_pr_iter finishes them (or write the partial of a row with cold sources
// to its first slot).  Sink mass goes to part[tile].
constexpr int kHotFill = kHotB / 4 / kHotBT;  // float4 loads per thread to fill one source block
static_assert(kHotFill * 4 * kHotBT == kHotB, "fill split");

// Block b's contributions (sources [b kHotB, min(hot, (b+1) kHotB))) into
// registers, float4 i of thread t at t + i kHotBT; past `hot`: zeros.
__device__ __forceinline__ void hot_fill_load(const float* __restrict__ contrib, int32_t hot, int32_t b,
                                              float4* f) {
    const int32_t src0 = b << kHotShift;
    const int32_t cnt = hot - src0 < kHotB ? hot - src0 : kHotB;
#pragma unroll
    for (int i = 0; i < kHotFill; ++i) {
        const int32_t q = static_cast<int32_t>(threadIdx.x) + i * kHotBT;  // float4 index
        if (4 * q + 4 <= cnt) {
            f[i] = reinterpret_cast<const float4*>(contrib + src0)[q];
        } else {
            f[i] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (4 * q + 0 < cnt) f[i].x = contrib[src0 + 4 * q + 0];
            if (4 * q + 1 < cnt) f[i].y = contrib[src0 + 4 * q + 1];
            if (4 * q + 2 < cnt) f[i].z = contrib[src0 + 4 * q + 2];
        }
    }
}

__device__ __forceinline__ void hot_fill_store(float* s_src, const float4* f) {
#pragma unroll
    for (int i = 0; i < kHotFill; ++i)
        reinterpret_cast<float4*>(s_src)[static_cast<int32_t>(threadIdx.x) + i * kHotBT] = f[i];
}

// The kHotK words of this thread in the chunk starting at p (zeros past end).
__device__ __forceinline__ void hot_words_load(const uint32_t* __restrict__ words, int32_t p, int32_t end,
                                               uint32_t* wn) {
    const int32_t my0 = p + static_cast<int32_t>(threadIdx.x) * kHotK;
#pragma unroll
    for (int k = 0; k < kHotK; ++k) wn[k] = my0 + k < end ? words[my0 + k] : 0u;
}

__device__ __forceinline__ unsigned long long hot_fix(double x) {
    return static_cast<unsigned long long>(__double2ll_rn(x * kHotFix));
}

__global__ void __launch_bounds__(kHotBT) k_pr_hot(
    const uint32_t* __restrict__ words, const int32_t* __restrict__ soff, const int32_t* __restrict__ tile_row,
    int32_t num_tiles, int32_t nb, int32_t hot, const int32_t* __restrict__ outdeg, const int32_t* __restrict__ slot,
    double* __restrict__ long_part, const float* __restrict__ contrib, float* __restrict__ next,
    const PRScalars* __restrict__ scal, double nd, double damping, double* __restrict__ part,
    double* __restrict__ rank_out, int last, int* tile_ctr) {
    __shared__ float s_src[kHotB];
    __shared__ unsigned long long s_acc[kHotT];
    __shared__ double s_red[kHotBT / 64];
    __shared__ int32_t s_tile;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const double base = scal->sink * damping / nd + (1.0 - damping) / nd;
    for (;;) {
        __syncthreads();  // s_tile, s_red and the accumulators are free again
        if (tid == 0) s_tile = atomicAdd(tile_ctr, 1);
        __syncthreads();
        const int32_t tile = s_tile;
        if (tile >= num_tiles) break;
        const int32_t row0 = tile_row[tile], rows = tile_row[tile + 1] - row0;  // rows <= kHotT
        for (int r = tid; r < rows; r += kHotBT) s_acc[r] = 0ULL;
        // Slices with at least kHotFillMin words read their sources from an
        // LDS copy of the block; smaller ones (the light rows' tiles) gather
        // from global memory, where one request per edge is cheaper than
        // filling 64 KB.  The next chunk's words are always in flight (the
        // tile's slices are contiguous, so the next chunk starts where this
        // one ends).
        const int32_t* ts = soff + int64_t(tile) * nb;
        const int32_t t_end = ts[nb];
        uint32_t wn[kHotK];
        hot_words_load(words, ts[0], t_end, wn);
        for (int32_t b = 0; b < nb; ++b) {
            const int32_t e0 = ts[b], e1 = ts[b + 1];
            const bool use_lds = e1 - e0 >= kHotFillMin;  // block-uniform
            const int32_t src0 = b << kHotShift;
            if (use_lds) {
                float4 f[kHotFill];
                hot_fill_load(contrib, hot, b, f);
                __syncthreads();  // the previous readers of s_src are done
                hot_fill_store(s_src, f);
                __syncthreads();
            }
            for (int32_t c0 = e0; c0 < e1; c0 += kHotChunk) {
                const int32_t c1 = e1 - c0 < kHotChunk ? e1 : c0 + kHotChunk;
                const int32_t my0 = c0 + tid * kHotK;
                const int32_t nvalid = c1 - my0 < 0 ? 0 : (c1 - my0 > kHotK ? kHotK : c1 - my0);
                uint32_t w[kHotK];
#pragma unroll
                for (int k = 0; k < kHotK; ++k) w[k] = wn[k];
                hot_words_load(words, c1, t_end, wn);  // the next chunk starts at c1
                float v[kHotK];
                if (use_lds) {
#pragma unroll
                    for (int k = 0; k < kHotK; ++k) v[k] = s_src[w[k] & (kHotB - 1)];
                } else {
#pragma unroll
                    for (int k = 0; k < kHotK; ++k)
                        v[k] = k < nvalid ? contrib[src0 + static_cast<int32_t>(w[k] & (kHotB - 1))] : 0.f;
                }
                // runs of equal rows (the words are sorted by row): each run's
                // FP64 sum goes to the row in fixed point, one atomic per run
                if (nvalid > 0) {
                    int32_t row = static_cast<int32_t>(w[0] >> kHotShift);
                    double run = static_cast<double>(v[0]);
#pragma unroll
                    for (int k = 1; k < kHotK; ++k) {
                        if (k < nvalid) {
                            const int32_t r = static_cast<int32_t>(w[k] >> kHotShift);
                            if (r != row) {
                                atomicAdd(&s_acc[row], hot_fix(run));
                                row = r;
                                run = 0.0;
                            }
                            run += static_cast<double>(v[k]);
                        }
                    }
                    atomicAdd(&s_acc[row], hot_fix(run));
                }
            }
        }
        __syncthreads();
        double sink = 0;  // per tile, whichever block takes it: part[tile] is deterministic
        for (int lr = tid; lr < rows; lr += kHotBT) {
            const int32_t row = row0 + lr;
            const double sum = static_cast<double>(s_acc[lr]) * (1.0 / kHotFix);
            const int32_t sl = slot[row];
            if (sl >= 0) {
                long_part[sl] = sum;
            } else {
                const double pr = base + damping * sum;
                const int32_t d = outdeg[row];
                if (d > 0) next[row] = static_cast<float>(pr / d);
                else sink += pr;
                if (last) rank_out[row] = pr;
            }
        }
        sink = wave_sum(sink);
        if (lane == 0) s_red[wave] = sink;
        __syncthreads();
        if (tid == 0) {
            double t = 0;
            for (int i = 0; i < kHotBT / 64; ++i) t += s_red[i];
            part[tile] = t;
        }
    }
}

// Pure gathers over a range of the layout's sources: each thread sums
// contrib[col[e]] over its edges with kGatherUnroll independent loads in
// flight and each wave stores one value.  No rows, no compensation, no LDS
// staging: the rate of this kernel is the ceiling of the per-edge gather
// pattern itself, against which the iteration's phases are measured.
constexpr int kGatherUnroll = 8;
__global__ void __launch_bounds__(kBT) k_gather_probe(const int32_t* __restrict__ col, int64_t e0, int64_t e1,
                                                      const float* __restrict__ contrib, float* __restrict__ out) {
    float acc = 0.f;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kBT * kGatherUnroll;
    for (int64_t b = e0 + static_cast<int64_t>(blockIdx.x) * kBT * kGatherUnroll + threadIdx.x; b < e1; b += stride) {
        int32_t idx[kGatherUnroll];
#pragma unroll
        for (int j = 0; j < kGatherUnroll; ++j) idx[j] = b + j * kBT < e1 ? col[b + j * kBT] : -1;
#pragma unroll
        for (int j = 0; j < kGatherUnroll; ++j) acc += idx[j] >= 0 ? contrib[idx[j]] : 0.f;
    }
    acc = wave_sum(acc);
    if (lane_id() == 0) out[(static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x) >> 6] = acc;
}

// Relabeled in-CSR: row r is old vertex perm[r], its in-neighbours renamed by inv.
__global__ void __launch_bounds__(kBT) k_permute_rows(const int32_t* __restrict__ perm, const int32_t* __restrict__ inv,
                                                      const int32_t* __restrict__ in_rowptr,
                                                      const int32_t* __restrict__ in_col,
                                                      const int32_t* __restrict__ rowptr_p, int32_t* __restrict__ col_p,
                                                      int32_t n) {
    const int lane = lane_id();
    const int64_t wave_global = (static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x) >> 6;
    const int64_t wave_stride = (static_cast<int64_t>(gridDim.x) * kBT) >> 6;
    for (int64_t r = wave_global; r < n; r += wave_stride) {
        const int32_t v = perm[r];
        const int32_t ob = in_rowptr[v], d = in_rowptr[v + 1] - ob, nb = rowptr_p[r];
        for (int32_t i = lane; i < d; i += 64) col_p[nb + i] = inv[in_col[ob + i]];
    }
}

// Relabeled sources below `hot` (the highest out-degrees) are "hot": their
// FP32 contributions stay in every XCD's L2.  A row's other sources are cut
// at multiples of `window` into cold segments; the hot parts of all rows are
// processed first, then the cold segments window by window, each writing a
// partial that k_pr_long adds to the row's hot partial.  Keeping cold
// gathers out of the hot phase stops them evicting hot lines.  Automatic
// split: none while the whole contribution array fits in 8 MB, else 2^19 hot
// sources (2 MB) and 2^23-id windows (a single cold window up to scale 24).
// Measured per iteration on R-MAT scale 24 (k_pr_iter + k_pr_long): no split
// 1839 us; hot 2^19 with windows 2^20 1754, 2^21 1719, 2^23 1697; hot 2^18
// or 2^20 are slower.
// code bits 16-17 choose the hot phase: 2 = source-blocked (k_pr_hot),
// anything else = per-edge gathers.  Code 0 keeps the gathers: on R-MAT
// s24 ef16 the source-blocked phase measured 1.26 ms against 1.09 ms for the
// gathers (it trades L2 requests for LDS fills, FP64 runs and 64-bit LDS
// atomics, and is instruction-bound), so it is available only on request.
void pr_split(int32_t n, int32_t code, int32_t* hot, int32_t* window, bool* blocked) {
    const int form = (code >> 16) & 3;
    if (form == 1 || form == 2) {
        *hot = static_cast<int32_t>(std::min<int64_t>(int64_t(1) << std::min(code & 0xFF, 30), n));
        *window = static_cast<int32_t>(int64_t(1) << std::min(std::max((code >> 8) & 0xFF, 1), 30));
        *blocked = ((code >> 16) & 3) == 2;
    } else if (n > (1 << 21)) {
        *hot = 1 << 19;
        *window = 1 << 23;
        *blocked = false;
    } else {
        *hot = n;
        *window = 1 << 20;
        *blocked = false;
    }
}

// Slice keys of the hot edges: tile * blocks + source block, and the packed
// word (row in tile << kHotShift | source in block).  One wave per row.
__global__ void __launch_bounds__(kBT) k_hot_keys(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                                  int32_t n, const int32_t* __restrict__ tile_row, int32_t tiles,
                                                  int32_t nb, uint32_t* __restrict__ key, uint32_t* __restrict__ word) {
    const int lane = lane_id();
    const int64_t wave_global = (static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x) >> 6;
    const int64_t wave_stride = (static_cast<int64_t>(gridDim.x) * kBT) >> 6;
    for (int64_t r = wave_global; r < n; r += wave_stride) {
        int32_t lo = 0, hi = tiles;  // the tile t with tile_row[t] <= r < tile_row[t + 1]
        while (hi - lo > 1) {
            const int32_t mid = (lo + hi) >> 1;
            if (tile_row[mid] <= r) lo = mid;
            else hi = mid;
        }
        const uint32_t tile_key = static_cast<uint32_t>(lo) * static_cast<uint32_t>(nb);
        const uint32_t lr = static_cast<uint32_t>(r - tile_row[lo]) << kHotShift;
        for (int32_t e = rowptr[r] + lane; e < rowptr[r + 1]; e += 64) {
            const uint32_t src = static_cast<uint32_t>(col[e]);
            key[e] = tile_key + (src >> kHotShift);
            word[e] = lr | (src & (kHotB - 1));
        }
    }
}

// soff[k] = first position of slice k in the sorted keys (empty slices get
// the position of the next non-empty one); soff[num_keys] = count.
__global__ void k_hot_soff(const uint32_t* __restrict__ key, int64_t count, uint32_t num_keys,
                           int32_t* __restrict__ soff) {
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e <= count;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t prev = e > 0 ? static_cast<int64_t>(key[e - 1]) : -1;
        const int64_t cur = e < count ? static_cast<int64_t>(key[e]) : static_cast<int64_t>(num_keys);
        for (int64_t k = prev + 1; k <= cur; ++k) soff[k] = static_cast<int32_t>(e);
    }
}

// Regroups the hot parts (rows [0, n) of the layout's rowptr / col, sources
// ascending in each row) into (row tile, source block) slices with one
// stable radix sort by slice key: inside a slice the rows stay in order and
// each row's sources ascending.
//
// Tiles are cut by rows (at most kHotT) and by hot edges (at most about
// 1/512 of them), so a few hub rows make a tile of their own and the atomic
// tile counter balances the blocks.
int build_hot_slices(PRLayout& L, int32_t n, int64_t num_hot, const std::vector<int32_t>& hot_len, hipStream_t s) {
    std::vector<int32_t> tile_row;
    {
        const int64_t target = std::max<int64_t>(1, num_hot / 512);
        int64_t rows = 0, edges = 0;
        for (int32_t r = 0; r < n; ++r) {
            if (rows == 0 || rows == kHotT || edges + hot_len[r] > target) {
                tile_row.push_back(r);
                rows = 0;
                edges = 0;
            }
            ++rows;
            edges += hot_len[r];
        }
        tile_row.push_back(n);
    }
    L.hot_tiles = static_cast<int32_t>(tile_row.size() - 1);
    L.hot_blocks = static_cast<int32_t>(std::max<int64_t>(1, ceil_div(L.hot, kHotB)));
    const int64_t num_keys = int64_t(L.hot_tiles) * L.hot_blocks;
    if (num_keys >= (int64_t(1) << 31) || num_hot >= (int64_t(1) << 31))
        return set_error(GHET_ERR_UNSUPPORTED, "PageRank hot slices: too many");
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t words_b = al(4 * size_t(std::max<int64_t>(num_hot, 1))), soff_b = al(4 * size_t(num_keys + 1));
    GHET_HIP(hipMalloc(&L.hot_mem, words_b + soff_b + al(4 * tile_row.size())));
    L.hot_words = static_cast<uint32_t*>(L.hot_mem);
    L.hot_soff = reinterpret_cast<int32_t*>(static_cast<char*>(L.hot_mem) + words_b);
    L.hot_tile_row = reinterpret_cast<int32_t*>(static_cast<char*>(L.hot_mem) + words_b + soff_b);
    GHET_HIP(hipMemcpyAsync(L.hot_tile_row, tile_row.data(), 4 * tile_row.size(), hipMemcpyHostToDevice, s));
    unsigned bits = 1;
    while ((int64_t(1) << bits) < num_keys) ++bits;
    uint32_t *ka = nullptr, *kb = nullptr, *wa = nullptr;
    const uint32_t* sorted_keys = nullptr;
    void* tmp = nullptr;
    size_t tb = 0;
    hipError_t e = hipSuccess;
    if (num_hot > 0) {
        e = hipMalloc(&ka, 4 * size_t(num_hot));
        if (e == hipSuccess) e = hipMalloc(&kb, 4 * size_t(num_hot));
        if (e == hipSuccess) e = hipMalloc(&wa, 4 * size_t(num_hot));
        if (e == hipSuccess) {
            const unsigned grid = static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, kBT / 64), 65535)));
            hipLaunchKernelGGL(k_hot_keys, dim3(grid), dim3(kBT), 0, s, L.rowptr, L.col, n, L.hot_tile_row, L.hot_tiles,
                               L.hot_blocks, ka, wa);
            e = hipGetLastError();
        }
        rocprim::double_buffer<uint32_t> kd(ka, kb);
        rocprim::double_buffer<uint32_t> vd(wa, L.hot_words);
        if (e == hipSuccess) e = rocprim::radix_sort_pairs(nullptr, tb, kd, vd, static_cast<unsigned>(num_hot), 0, bits, s);
        if (e == hipSuccess) e = hipMalloc(&tmp, std::max<size_t>(tb, 256));
        if (e == hipSuccess) e = rocprim::radix_sort_pairs(tmp, tb, kd, vd, static_cast<unsigned>(num_hot), 0, bits, s);
        if (e == hipSuccess && vd.current() != L.hot_words)
            e = hipMemcpyAsync(L.hot_words, vd.current(), 4 * size_t(num_hot), hipMemcpyDeviceToDevice, s);
        sorted_keys = kd.current();
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_hot_soff, dim3(static_cast<unsigned>(std::min<int64_t>(ceil_div(num_hot + 1, 256), 65535))),
                               dim3(256), 0, s, sorted_keys, num_hot, static_cast<uint32_t>(num_keys), L.hot_soff);
            e = hipGetLastError();
        }
    } else {
        hipLaunchKernelGGL(k_hot_soff, dim3(1), dim3(256), 0, s, static_cast<const uint32_t*>(nullptr), int64_t(0),
                           static_cast<uint32_t>(num_keys), L.hot_soff);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    for (void* q : {static_cast<void*>(ka), static_cast<void*>(kb), static_cast<void*>(wa), tmp})
        if (q) hipFree(q);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return set_error(GHET_ERR_HIP, "PageRank hot slices: %s", hipGetErrorString(e));
    }
    return GHET_OK;
}

// ---------------------------------------------------------------------------
// Layout construction on the device.  The relabeled in-CSR is permuted and
// every row's sources sorted by a segmented radix sort; hot lengths, cold
// segments, slots and the combined offsets then come from flags and scans
// over edges, segments and rows.  Only the tile packing (a greedy walk over
// the combined offsets) runs on the host.

__device__ __forceinline__ int32_t pr_nslots(int32_t len) {
    return len > kTileNnz ? static_cast<int32_t>((len + kTileNnz - 1) / kTileNnz) : 1;
}

// Class of relabeled source u: 0 hot (u < hot), else 1 + its window.
__device__ __forceinline__ uint32_t pr_class(int32_t u, int32_t hot, int wshift) {
    return u < hot ? 0u : 1u + (static_cast<uint32_t>(u - hot) >> wshift);
}

// Out-degree order: key ~outdeg (ascending = out-degree descending), value v;
// a stable radix sort then keeps equal degrees in id order.
__global__ void __launch_bounds__(kBT) k_pr_deg_keys(const int32_t* __restrict__ out_rp, int32_t n,
                                                     uint32_t* __restrict__ key, int32_t* __restrict__ val) {
    for (int64_t v = static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x; v < n; v += static_cast<int64_t>(gridDim.x) * kBT) {
        key[v] = ~static_cast<uint32_t>(out_rp[v + 1] - out_rp[v]);
        val[v] = static_cast<int32_t>(v);
    }
}

// inv[perm[r]] = r; in- and out-degree of relabeled row r.
__global__ void __launch_bounds__(kBT) k_pr_relabel(const int32_t* __restrict__ perm, int32_t n,
                                                    const int32_t* __restrict__ in_rp, const int32_t* __restrict__ out_rp,
                                                    int32_t* __restrict__ inv, int32_t* __restrict__ indeg,
                                                    int32_t* __restrict__ outdeg) {
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x; r < n; r += static_cast<int64_t>(gridDim.x) * kBT) {
        const int32_t v = perm[r];
        inv[v] = static_cast<int32_t>(r);
        indeg[r] = in_rp[v + 1] - in_rp[v];
        outdeg[r] = out_rp[v + 1] - out_rp[v];
    }
}

// Per row: hot length (its sources are ascending, so the first source at
// or above `hot` ends the hot part), and a mark at the row's first edge.
__global__ void __launch_bounds__(kBT) k_pr_hot_len(const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                                    int32_t n, int32_t hot, int32_t* __restrict__ hot_len,
                                                    uint8_t* __restrict__ row_start) {
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x; r < n; r += static_cast<int64_t>(gridDim.x) * kBT) {
        int32_t lo = rp[r], hi = rp[r + 1];
        if (lo < hi) row_start[lo] = 1;
        const int32_t b = lo;
        while (lo < hi) {
            const int32_t mid = lo + ((hi - lo) >> 1);
            if (col[mid] < hot) lo = mid + 1;
            else hi = mid;
        }
        hot_len[r] = lo - b;
    }
}

// Edge e opens a cold segment when its source is cold and it starts its row
// or follows a source of another class.
__global__ void __launch_bounds__(kBT) k_pr_seg_flags(const int32_t* __restrict__ col, const uint8_t* __restrict__ row_start,
                                                      int64_t m, int32_t hot, int wshift, int32_t* __restrict__ flag) {
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x; e < m; e += static_cast<int64_t>(gridDim.x) * kBT) {
        const uint32_t c = pr_class(col[e], hot, wshift);
        flag[e] = (c > 0 && (row_start[e] || pr_class(col[e - (e > 0)], hot, wshift) != c || e == 0)) ? 1 : 0;
    }
}

__global__ void __launch_bounds__(kBT) k_pr_seg_start(const int32_t* __restrict__ flag, const int32_t* __restrict__ pos,
                                                      int64_t m, int32_t* __restrict__ seg_start) {
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x; e < m; e += static_cast<int64_t>(gridDim.x) * kBT)
        if (flag[e]) seg_start[pos[e]] = static_cast<int32_t>(e);
}

// Per segment q (row-major order): its row, length, window key and slot
// count.
__global__ void __launch_bounds__(kBT) k_pr_seg_info(const int32_t* __restrict__ seg_start, int32_t S, int64_t m,
                                                     const int32_t* __restrict__ rp, int32_t n,
                                                     const int32_t* __restrict__ col, int32_t hot, int wshift,
                                                     int32_t* __restrict__ seg_row, int32_t* __restrict__ seg_len,
                                                     uint32_t* __restrict__ seg_key, int32_t* __restrict__ seg_ns) {
    for (int64_t q = static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x; q < S; q += static_cast<int64_t>(gridDim.x) * kBT) {
        const int32_t e = seg_start[q];
        int32_t lo = 0, hi = n;  // the row r with rp[r] <= e < rp[r + 1]: last r with rp[r] <= e
        while (hi - lo > 1) {
            const int32_t mid = lo + ((hi - lo) >> 1);
            if (rp[mid] <= e) lo = mid;
            else hi = mid;
        }
        const int64_t next = q + 1 < S ? seg_start[q + 1] : m;
        const int32_t end = static_cast<int32_t>(next < rp[lo + 1] ? next : rp[lo + 1]);
        const uint32_t key = pr_class(col[e], hot, wshift) - 1;
        seg_row[q] = lo;
        seg_len[q] = end - e;
        seg_key[q] = key;
        seg_ns[q] = pr_nslots(end - e);
    }
}

// Per row: partial slots (hot part and cold segments), and the offsets'
// inputs for the rows finished by k_pr_long.
__global__ void __launch_bounds__(kBT) k_pr_row_slots(const int32_t* __restrict__ rp, const int32_t* __restrict__ pos,
                                                      const int32_t* __restrict__ sn, const int32_t* __restrict__ hot_len,
                                                      int32_t n, int blocked, int32_t* __restrict__ nslots,
                                                      int32_t* __restrict__ partial) {
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x; r < n; r += static_cast<int64_t>(gridDim.x) * kBT) {
        const int32_t cold = sn[pos[rp[r + 1]]] - sn[pos[rp[r]]];
        const int32_t hl = hot_len[r];
        const int32_t k = cold > 0 ? (blocked ? 1 : pr_nslots(hl)) + cold : ((!blocked && hl > kTileNnz) ? pr_nslots(hl) : 0);
        nslots[r] = k;
        partial[r] = k > 0 ? 1 : 0;
    }
}

// Hot parts: slot of each row, its row of k_pr_long, and its hot edges
// copied to the combined column array (one wave per row).
__global__ void __launch_bounds__(kBT) k_pr_fill_hot(const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                                     const int32_t* __restrict__ hot_len, const int32_t* __restrict__ rp_all,
                                                     const int32_t* __restrict__ nslots, const int32_t* __restrict__ slot_base,
                                                     const int32_t* __restrict__ lpos, int32_t n,
                                                     int32_t* __restrict__ slot_all, int4* __restrict__ long_rows,
                                                     int32_t* __restrict__ col_all) {
    const int lane = lane_id();
    const int64_t wave_global = (static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x) >> 6;
    const int64_t wave_stride = (static_cast<int64_t>(gridDim.x) * kBT) >> 6;
    for (int64_t r = wave_global; r < n; r += wave_stride) {
        const int32_t k = nslots[r], sb = slot_base[r];
        if (lane == 0) {
            slot_all[r] = k > 0 ? sb : -1;
            if (k > 0) long_rows[lpos[r]] = make_int4(static_cast<int32_t>(r), sb, k, 0);
        }
        const int32_t b = rp[r], d = hot_len[r], o = rp_all[r];
        for (int32_t i = lane; i < d; i += 64) col_all[o + i] = col[b + i];
    }
}

// Cold segments in window-major order: segment sid is row-major segment
// order[sid]; its offset, slot (after the row's hot slots and the slots of
// the row's earlier segments) and edges.  Segments up to kSegLight edges
// are copied lane-serially, longer ones by the whole wave.
constexpr int kSegLight = 16;
__global__ void __launch_bounds__(kBT) k_pr_fill_cold(
    const int32_t* __restrict__ order, int32_t S, const int32_t* __restrict__ seg_start,
    const int32_t* __restrict__ seg_row, const int32_t* __restrict__ coff, const int32_t* __restrict__ sn,
    const int32_t* __restrict__ pos, const int32_t* __restrict__ rp, const int32_t* __restrict__ hot_len,
    const int32_t* __restrict__ slot_base, const int32_t* __restrict__ col, int32_t n, int64_t num_hot, int blocked,
    int32_t* __restrict__ rp_all, int32_t* __restrict__ slot_all, int32_t* __restrict__ col_all) {
    const int lane = lane_id();
    const int64_t wave_global = (static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x) >> 6;
    const int64_t wave_stride = (static_cast<int64_t>(gridDim.x) * kBT) >> 6;
    for (int64_t g0 = wave_global * 64; g0 < S; g0 += wave_stride * 64) {
        const int64_t sid = g0 + lane;
        int32_t b = 0, len = 0, o = 0;
        if (sid < S) {
            const int32_t q = order[sid], r = seg_row[q];
            b = seg_start[q];
            o = static_cast<int32_t>(num_hot + coff[sid]);
            len = coff[sid + 1] - coff[sid];
            const int32_t hs = blocked ? 1 : pr_nslots(hot_len[r]);
            rp_all[n + sid] = o;
            slot_all[n + sid] = slot_base[r] + hs + sn[q] - sn[pos[rp[r]]];
        }
        if (len <= kSegLight)
            for (int32_t i = 0; i < len; ++i) col_all[o + i] = col[b + i];
        uint64_t heavy = __ballot(len > kSegLight);
        while (heavy) {
            const int srcl = __ffsll(static_cast<unsigned long long>(heavy)) - 1;
            heavy &= heavy - 1;
            const int32_t hb = __shfl(b, srcl), hl = __shfl(len, srcl), ho = __shfl(o, srcl);
            for (int32_t i = lane; i < hl; i += 64) col_all[ho + i] = col[hb + i];
        }
    }
}

__global__ void __launch_bounds__(kBT) k_pr_gather_len(const int32_t* __restrict__ order, const int32_t* __restrict__ seg_len,
                                                       int32_t S, int32_t* __restrict__ out) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x; i < S; i += static_cast<int64_t>(gridDim.x) * kBT)
        out[i] = seg_len[order[i]];
}

__global__ void __launch_bounds__(kBT) k_pr_iota(int32_t* __restrict__ out, int32_t count) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x; i < count; i += static_cast<int64_t>(gridDim.x) * kBT)
        out[i] = static_cast<int32_t>(i);
}

unsigned pr_grid(int64_t threads) {
    return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(ceil_div(threads, kBT), 65536)));
}

int build_pr_layout(ghet_graph* g, hipStream_t s, int32_t split_code) {
    PRLayout& L = g->pr;
    const int32_t n = g->n;
    int32_t H = 0, W = 0;
    bool blocked = false;
    pr_split(n, split_code, &H, &W, &blocked);
    if (L.built && L.hot == H && L.window == W && L.blocked == blocked) return GHET_OK;
    if (L.mem || L.hot_mem || L.items_mem) {
        GHET_HIP(hipStreamSynchronize(s));
        if (L.mem) GHET_HIP(hipFree(L.mem));
        if (L.hot_mem) GHET_HIP(hipFree(L.hot_mem));
        if (L.items_mem) GHET_HIP(hipFree(L.items_mem));
        L = PRLayout{};
    }
    const int64_t m = g->in.m;
    if (m >= INT32_MAX) return set_error(GHET_ERR_UNSUPPORTED, "PageRank layout: %lld edges", static_cast<long long>(m));
    int wshift = 0;
    while ((int64_t(1) << wshift) < W) ++wshift;
    const int32_t NW = n > H ? static_cast<int32_t>(ceil_div(n - H, W)) : 0;
    const unsigned gn = pr_grid(n), gm = pr_grid(m), gw = pr_grid(int64_t(n) * 64);
    DevTemp t;
    auto hip_ok = [&](hipError_t e) {
        if (t.err == hipSuccess) t.err = e;
        return t.err == hipSuccess;
    };
    auto fail = [&](const char* what) {
        (void)hipGetLastError();
        return set_error(GHET_ERR_HIP, "PageRank layout (%s): %s", what, hipGetErrorString(t.err));
    };
    const int64_t scan_max = std::max<int64_t>(m, n) + 1;
    int32_t* aux = t.get<int32_t>(scan_aux_ints(scan_max));
    auto scan = [&](const int32_t* in, int32_t* out, int64_t count) {
        if (t.err == hipSuccess && exclusive_scan_i32(in, out, count, aux, s) != GHET_OK) t.err = hipErrorUnknown;
    };
    void* sort_tmp = nullptr;
    size_t sort_bytes = 0;
    auto sort_space = [&](size_t need) {  // grows one scratch buffer for the rocPRIM sorts
        if (t.err != hipSuccess || need <= sort_bytes) return;
        sort_tmp = t.get<char>(need);
        sort_bytes = need;
    };

    // 1. Relabeling by out-degree, descending, stable by id.
    uint32_t* ka = t.get<uint32_t>(n);
    uint32_t* kb = t.get<uint32_t>(n);
    int32_t* va = t.get<int32_t>(n);
    int32_t* vb = t.get<int32_t>(n);
    int32_t* inv = t.get<int32_t>(n);
    int32_t* indeg = t.get<int32_t>(n);
    int32_t* outdeg = t.get<int32_t>(n);
    int32_t* rp = t.get<int32_t>(size_t(n) + 1);
    if (t.err != hipSuccess) return fail("alloc");
    hipLaunchKernelGGL(k_pr_deg_keys, dim3(gn), dim3(kBT), 0, s, g->out.rowptr, n, ka, va);
    rocprim::double_buffer<uint32_t> kd(ka, kb);
    rocprim::double_buffer<int32_t> vd(va, vb);
    size_t need = 0;
    hip_ok(rocprim::radix_sort_pairs(nullptr, need, kd, vd, static_cast<unsigned>(n), 0, 32, s));
    sort_space(need);
    hip_ok(rocprim::radix_sort_pairs(sort_tmp, need, kd, vd, static_cast<unsigned>(n), 0, 32, s));
    int32_t* perm = vd.current();
    hipLaunchKernelGGL(k_pr_relabel, dim3(gn), dim3(kBT), 0, s, perm, n, g->in.rowptr, g->out.rowptr, inv, indeg, outdeg);
    hip_ok(hipGetLastError());
    scan(indeg, rp, n);
    // 2. Relabeled in-lists, then every row's sources ascending.
    int32_t* pcol = t.get<int32_t>(size_t(m));
    int32_t* col = t.get<int32_t>(size_t(m));
    if (t.err != hipSuccess) return fail("relabel");
    hipLaunchKernelGGL(k_permute_rows, dim3(gw), dim3(kBT), 0, s, perm, inv, g->in.rowptr, g->in.col, rp, pcol, n);
    unsigned bits = 1;
    while ((int64_t(1) << bits) < n) ++bits;
    if (m > 0) {
        hip_ok(rocprim::segmented_radix_sort_keys(nullptr, need, pcol, col, static_cast<unsigned>(m), static_cast<unsigned>(n),
                                                  rp, rp + 1, 0, bits, s));
        sort_space(need);
        hip_ok(rocprim::segmented_radix_sort_keys(sort_tmp, need, pcol, col, static_cast<unsigned>(m),
                                                  static_cast<unsigned>(n), rp, rp + 1, 0, bits, s));
    }
    // 3. Hot lengths and cold segments (row-major numbering q).
    int32_t* hot_len = t.get<int32_t>(n);
    uint8_t* row_start = t.get<uint8_t>(size_t(m));
    int32_t* flag = t.get<int32_t>(size_t(m));
    int32_t* pos = t.get<int32_t>(size_t(m) + 1);
    int32_t* win_start = t.get<int32_t>(size_t(NW) + 1);
    if (t.err != hipSuccess) return fail("segments");
    hip_ok(hipMemsetAsync(row_start, 0, std::max<size_t>(size_t(m), 1), s));
    hipLaunchKernelGGL(k_pr_hot_len, dim3(gn), dim3(kBT), 0, s, rp, col, n, H, hot_len, row_start);
    hipLaunchKernelGGL(k_pr_seg_flags, dim3(gm), dim3(kBT), 0, s, col, row_start, m, H, wshift, flag);
    hip_ok(hipGetLastError());
    scan(flag, pos, m);
    int32_t S = 0;
    hip_ok(hipMemcpyAsync(&S, pos + m, 4, hipMemcpyDeviceToHost, s));
    hip_ok(hipStreamSynchronize(s));
    if (t.err != hipSuccess) return fail("segments");
    if (int64_t(n) + S >= INT32_MAX) return set_error(GHET_ERR_UNSUPPORTED, "PageRank layout: too many segments");
    int32_t* seg_start = t.get<int32_t>(S);
    int32_t* seg_row = t.get<int32_t>(S);
    int32_t* seg_len = t.get<int32_t>(S);
    int32_t* seg_ns = t.get<int32_t>(S);
    int32_t* sn = t.get<int32_t>(size_t(S) + 1);
    uint32_t* sk_a = t.get<uint32_t>(S);
    uint32_t* sk_b = t.get<uint32_t>(S);
    int32_t* so_a = t.get<int32_t>(S);
    int32_t* so_b = t.get<int32_t>(S);
    int32_t* clen = t.get<int32_t>(S);
    int32_t* coff = t.get<int32_t>(size_t(S) + 1);
    if (t.err != hipSuccess) return fail("segments");
    const unsigned gs = pr_grid(S);
    hipLaunchKernelGGL(k_pr_seg_start, dim3(gm), dim3(kBT), 0, s, flag, pos, m, seg_start);
    hipLaunchKernelGGL(k_pr_seg_info, dim3(gs), dim3(kBT), 0, s, seg_start, S, m, rp, n, col, H, wshift, seg_row, seg_len,
                       sk_a, seg_ns);
    hip_ok(hipGetLastError());
    scan(seg_ns, sn, S);
    // 4. Slots per row and their offsets; the rows k_pr_long finishes.
    int32_t* nslots = t.get<int32_t>(n);
    int32_t* slot_base = t.get<int32_t>(size_t(n) + 1);
    int32_t* partial = t.get<int32_t>(n);
    int32_t* lpos = t.get<int32_t>(size_t(n) + 1);
    if (t.err != hipSuccess) return fail("slots");
    hipLaunchKernelGGL(k_pr_row_slots, dim3(gn), dim3(kBT), 0, s, rp, pos, sn, hot_len, n, blocked ? 1 : 0, nslots, partial);
    hip_ok(hipGetLastError());
    scan(nslots, slot_base, n);
    scan(partial, lpos, n);
    int32_t counts[2] = {0, 0};  // slots, rows with partials
    hip_ok(hipMemcpyAsync(counts, slot_base + n, 4, hipMemcpyDeviceToHost, s));
    hip_ok(hipMemcpyAsync(counts + 1, lpos + n, 4, hipMemcpyDeviceToHost, s));
    // 5. Window-major segment order: a stable radix sort of the window keys.
    rocprim::double_buffer<uint32_t> skd(sk_a, sk_b);
    rocprim::double_buffer<int32_t> sod(so_a, so_b);
    hipLaunchKernelGGL(k_pr_iota, dim3(gs), dim3(kBT), 0, s, so_a, S);
    if (NW > 1 && S > 0) {
        unsigned wb = 1;
        while ((int64_t(1) << wb) < NW) ++wb;
        hip_ok(rocprim::radix_sort_pairs(nullptr, need, skd, sod, static_cast<unsigned>(S), 0, wb, s));
        sort_space(need);
        hip_ok(rocprim::radix_sort_pairs(sort_tmp, need, skd, sod, static_cast<unsigned>(S), 0, wb, s));
    }
    const int32_t* order = sod.current();
    // first segment of each window in that order (win_start[NW] = S)
    hipLaunchKernelGGL(k_hot_soff, dim3(pr_grid(int64_t(S) + 1)), dim3(kBT), 0, s, skd.current(), int64_t(S),
                       static_cast<uint32_t>(NW), win_start);
    hipLaunchKernelGGL(k_pr_gather_len, dim3(gs), dim3(kBT), 0, s, order, seg_len, S, clen);
    hip_ok(hipGetLastError());
    scan(clen, coff, S);
    std::vector<int32_t> wstart(static_cast<size_t>(NW) + 1);
    hip_ok(hipMemcpyAsync(wstart.data(), win_start, 4 * wstart.size(), hipMemcpyDeviceToHost, s));
    hip_ok(hipStreamSynchronize(s));
    if (t.err != hipSuccess) return fail("slots");
    const int32_t num_slots = counts[0], num_long_rows = counts[1];
    // 6. The layout arrays and the combined offsets, slots and columns.
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t rows_all = static_cast<size_t>(n) + S;
    const size_t bytes = al(sizeof(int4) * size_t(num_long_rows)) + 2 * al(4 * size_t(n)) + al(4 * (rows_all + 1)) +
                         al(4 * std::max<size_t>(rows_all, 1)) + al(4 * std::max<size_t>(size_t(m), 1)) + 256;
    GHET_HIP(hipMalloc(&L.mem, bytes));
    char* p = static_cast<char*>(L.mem);
    L.long_rows = p;
    p += al(sizeof(int4) * size_t(num_long_rows));
    auto take = [&](size_t cnt) { int32_t* q = reinterpret_cast<int32_t*>(p); p += al(4 * cnt); return q; };
    L.outdeg = take(size_t(n));
    L.perm = take(size_t(n));
    L.rowptr = take(rows_all + 1);
    L.slot = take(std::max<size_t>(rows_all, 1));
    L.col = take(std::max<size_t>(size_t(m), 1));
    // k_pr_iter's masked lanes gather contrib[col[0]]: col[0] must be a
    // vertex id, also when there are no edges
    if (m == 0) GHET_CHECK(fill<int32_t>(L.col, 0, 1, s));
    hip_ok(hipMemcpyAsync(L.outdeg, outdeg, 4 * size_t(n), hipMemcpyDeviceToDevice, s));
    hip_ok(hipMemcpyAsync(L.perm, perm, 4 * size_t(n), hipMemcpyDeviceToDevice, s));
    scan(hot_len, L.rowptr, n);  // hot parts' offsets; L.rowptr[n] = hot edges
    int32_t num_hot32 = 0;
    hip_ok(hipMemcpyAsync(&num_hot32, L.rowptr + n, 4, hipMemcpyDeviceToHost, s));
    hip_ok(hipStreamSynchronize(s));
    if (t.err != hipSuccess) return fail("fill");
    const int64_t num_hot = num_hot32;
    hipLaunchKernelGGL(k_pr_fill_hot, dim3(gw), dim3(kBT), 0, s, rp, col, hot_len, L.rowptr, nslots, slot_base, lpos, n,
                       L.slot, static_cast<int4*>(L.long_rows), L.col);
    hipLaunchKernelGGL(k_pr_fill_cold, dim3(pr_grid(int64_t(ceil_div(S, 64)) * 64)), dim3(kBT), 0, s, order, S, seg_start,
                       seg_row, coff, sn, pos, rp, hot_len, slot_base, col, n, num_hot, blocked ? 1 : 0, L.rowptr, L.slot,
                       L.col);
    hip_ok(hipGetLastError());
    if (t.err != hipSuccess) return fail("fill");
    GHET_CHECK(fill<int32_t>(L.rowptr + rows_all, static_cast<int32_t>(m), 1, s));
    // 7. Work items over the combined row space [0, n+S) on the host: chunks
    // of long rows or segments (more than kTileNnz edges) write their slots;
    // tiles pack consecutive short ones.  Hot items first, then one group per
    // window.
    std::vector<int32_t> rp_all(rows_all + 1), slot_all(std::max<size_t>(rows_all, 1));
    hip_ok(hipMemcpyAsync(rp_all.data(), L.rowptr, 4 * (rows_all + 1), hipMemcpyDeviceToHost, s));
    hip_ok(hipMemcpyAsync(slot_all.data(), L.slot, 4 * rows_all, hipMemcpyDeviceToHost, s));
    std::vector<int32_t> hot_len_h;
    if (blocked) {
        hot_len_h.resize(static_cast<size_t>(n));
        hip_ok(hipMemcpyAsync(hot_len_h.data(), hot_len, 4 * size_t(n), hipMemcpyDeviceToHost, s));
    }
    hip_ok(hipStreamSynchronize(s));
    if (t.err != hipSuccess) return fail("download");
    std::vector<PRItem> items;
    auto emit = [&](int32_t lo, int32_t hi) {  // rows [lo, hi) of the combined space
        for (int32_t r = lo; r < hi;) {
            const int32_t d = rp_all[r + 1] - rp_all[r];
            if (d > kTileNnz) {
                const int32_t nch = static_cast<int32_t>(ceil_div(d, kTileNnz));
                for (int32_t k = 0; k < nch; ++k) {
                    PRItem it{};
                    it.row = r;
                    it.row_end = -1;
                    it.e0 = rp_all[r] + k * kTileNnz;
                    it.e1 = std::min(rp_all[r + 1], it.e0 + kTileNnz);
                    it.chunk = slot_all[r] + k;
                    items.push_back(it);
                }
                ++r;
                continue;
            }
            const int32_t first = r;
            int32_t nnz = 0;
            while (r < hi && r - first < kTileRows) {
                const int32_t dr = rp_all[r + 1] - rp_all[r];
                if (dr > kTileNnz || nnz + dr > kTileNnz) break;
                nnz += dr;
                ++r;
            }
            PRItem it{};
            it.row = first;
            it.row_end = r;
            it.e0 = rp_all[first];
            it.e1 = rp_all[r];
            items.push_back(it);
        }
    };
    if (!blocked) emit(0, n);  // source-blocked: k_pr_hot covers the hot parts
    const int32_t num_hot_items = static_cast<int32_t>(items.size());
    if (wstart[0] != 0 || wstart[NW] != S) return set_error(GHET_ERR_INVALID, "PageRank layout: window starts do not cover the segments");
    for (int32_t w = 0; w < NW; ++w) {
        if (wstart[w] > wstart[w + 1]) return set_error(GHET_ERR_INVALID, "PageRank layout: window starts out of order");
        emit(n + wstart[w], n + wstart[w + 1]);
    }
    int32_t chunks = 0, tiles = 0;
    for (const PRItem& it : items) (it.row_end < 0 ? chunks : tiles) += 1;
    // Every index a launch will use, checked once here: offsets and slots of
    // the combined rows, item edge ranges, rows and chunk slots.
    {
        const int64_t ra = static_cast<int64_t>(rows_all);
        std::atomic<bool> ok{rp_all[0] == 0 && rp_all[ra] == m && num_hot >= 0 && num_hot <= m};
        parallel_for(ra, hw_threads(0), [&](int64_t b, int64_t e, int) {
            bool good = true;
            for (int64_t r = b; good && r < e; ++r)
                good = rp_all[r] <= rp_all[r + 1] && slot_all[r] >= -1 && slot_all[r] < num_slots &&
                       (r < n || slot_all[r] >= 0);
            if (!good) ok = false;
        });
        for (const PRItem& it : items) {
            if (!ok) break;
            ok = it.e0 >= 0 && it.e0 <= it.e1 && it.e1 <= m && it.row >= 0 && it.row < ra &&
                 (it.row_end < 0 ? (it.chunk >= 0 && it.chunk < num_slots && it.e1 - it.e0 <= kTileNnz)
                                 : (it.row_end > it.row && it.row_end <= ra && it.row_end - it.row <= kTileRows &&
                                    it.e0 == rp_all[it.row] && it.e1 == rp_all[it.row_end] && it.e1 - it.e0 <= kTileNnz));
        }
        if (!ok) return set_error(GHET_ERR_INVALID, "PageRank layout failed its bounds check");
    }
    GHET_HIP(hipMalloc(&L.items_mem, std::max<size_t>(sizeof(PRItem) * items.size(), 256)));
    L.items = L.items_mem;
    if (!items.empty())
        GHET_HIP(hipMemcpyAsync(L.items, items.data(), sizeof(PRItem) * items.size(), hipMemcpyHostToDevice, s));
    GHET_HIP(hipStreamSynchronize(s));
    L.num_tiles = tiles;
    L.num_long_chunks = chunks;
    L.num_long_rows = num_long_rows;
    L.num_slots = num_slots;
    L.num_segments = S;
    L.num_items = static_cast<int32_t>(items.size());
    L.num_hot_items = num_hot_items;
    L.num_hot_edges = num_hot;
    L.hot = H;
    L.window = W;
    L.blocked = blocked;
    if (blocked) GHET_CHECK(build_hot_slices(L, n, num_hot, hot_len_h, s));
    L.built = true;
    return GHET_OK;
}

}  // namespace

int run_pr(ghet_graph* g, int32_t iters, double damping, double* pr_out, const ghet_opts* o, ghet_run_stats* st) {
    const int32_t n = g->n;
    if (iters < 0) return set_error(GHET_ERR_INVALID, "iterations %d < 0", iters);
    if (!g->in.present() || g->h_in_rowptr.empty()) return set_error(GHET_ERR_INVALID, "PageRank needs the in-CSR");
    if (n == 0) return GHET_OK;
    hipStream_t s = pick_stream(g, o);
    const bool normalise = !(o && o->size >= sizeof(ghet_opts) && (o->reserved[0] & 1));
    const int32_t split_code = (o && o->size >= sizeof(ghet_opts)) ? o->reserved[4] : 0;
    // Form (code bits 16-17): 3 = destination bins (pr_bin.hip); 1, 2 = this
    // file's gathers with the given split; 0 = automatic: bins above 2^21
    // vertices unless bit 18 asks for the gathers.
    const int form = (split_code >> 16) & 3;
    if (form == 3) return run_pr_binned(g, iters, damping, pr_out, normalise, o, st);
    if (form == 0 && !((split_code >> 18) & 1) && !g->prb.unfit && pr_binned_default(n, g->in.m)) {
        const int rc = run_pr_binned(g, iters, damping, pr_out, normalise, o, st);
        if (rc != GHET_ERR_UNSUPPORTED) return rc;
        // more bins than the binned form's tables hold: this graph takes the
        // gathers, now and on later calls
        g->prb.unfit = true;
        clear_error();
    }
    GHET_CHECK(build_pr_layout(g, s, split_code));
    PRLayout& L = g->pr;
    // One persistent k_pr_iter launch per iteration walks all items (hot
    // parts, then the cold segments) with as many blocks as stay resident,
    // and k_pr_long finishes the rows with partials.  The source-blocked
    // phase (k_pr_hot) runs before it over the hot parts.  With opts
    // reserved[0] bit 1 the gathers run as two launches, hot items then
    // cold items, with a HIP event between them: the phases are timed
    // separately (the split costs about 4 % at scale 24: the hot launch's
    // tail no longer overlaps the cold items).
    int occ = 0, occ_hot = 0, cus = 0;
    GHET_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(k_pr_iter), kBT, 0));
    GHET_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_hot, reinterpret_cast<const void*>(k_pr_hot), kHotBT, 0));
    GHET_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device));
    auto resident = [&](int o, int64_t work) {
        return static_cast<int32_t>(std::max<int64_t>(0, std::min<int64_t>(int64_t(std::max(o, 1)) * std::max(cus, 1), work)));
    };
    const bool phases = (o && o->size >= sizeof(ghet_opts) && (o->reserved[0] & 2)) || L.blocked;
    const int32_t hot_items = phases ? L.num_hot_items : 0, cold_items = L.num_items - hot_items;
    const int32_t hot_blocks = L.blocked ? std::max(1, resident(occ_hot, L.hot_tiles)) : resident(occ, hot_items);
    const int32_t cold_blocks = resident(occ, cold_items);
    const int32_t init_blocks = static_cast<int32_t>(std::min<int64_t>(ceil_div(n, kBT), 4096));
    // one row per thread: k_pr_long's loads (row, slots, out-degree) are a
    // dependent chain, so every row gets its own thread (1024 blocks with a
    // grid-stride loop took 70 us at scale 24)
    const int32_t long_blocks = static_cast<int32_t>(std::max<int64_t>(1, std::min<int64_t>(ceil_div(L.num_long_rows, kBT), 65535)));
    // sink partials: the hot phase's (one per tile for k_pr_hot, one per
    // block for k_pr_iter), the cold phase's, then k_pr_long's
    const int32_t hot_parts = L.blocked ? L.hot_tiles : hot_blocks;
    const int32_t parts = hot_parts + cold_blocks + long_blocks;
    void* p;
    float *ca, *cb;
    double *rank, *out, *part, *long_part;
    PRScalars* scal;
    GHET_CHECK(g->ws.get(WS_PR_CONTRIB, 4 * size_t(n), &p)); ca = static_cast<float*>(p);
    GHET_CHECK(g->ws.get(WS_LABEL_A, 4 * size_t(n), &p)); cb = static_cast<float*>(p);
    GHET_CHECK(g->ws.get(WS_PR_RANK, 8 * size_t(n), &p)); rank = static_cast<double*>(p);
    GHET_CHECK(g->ws.get(WS_PR_OUT, 8 * size_t(n), &p)); out = static_cast<double*>(p);
    GHET_CHECK(g->ws.get(WS_PR_PART, 8 * size_t(std::max(parts, init_blocks) + 1), &p)); part = static_cast<double*>(p);
    GHET_CHECK(g->ws.get(WS_PR_LONGPART, 8 * size_t(L.num_slots + 1), &p)); long_part = static_cast<double*>(p);
    GHET_CHECK(g->ws.get(WS_COUNTERS, 256, &p)); scal = static_cast<PRScalars*>(p);
    int* tile_ctr = reinterpret_cast<int*>(static_cast<char*>(p) + 128);
    const double inv_n = 1.0 / n;
    const bool timing = st != nullptr;
    double iter_ms = 0, hot_ms = 0;
    hipEvent_t* kev = nullptr;
    if (timing) GHET_CHECK(kernel_events(g, 3 * iters, &kev));

    GHET_HIP(hipEventRecord(g->ev_begin, s));
    hipLaunchKernelGGL(k_pr_init, dim3(init_blocks), dim3(kBT), 0, s, L.outdeg, n, inv_n, ca, part);
    hipLaunchKernelGGL(k_pr_reduce, dim3(1), dim3(1024), 0, s, part, init_blocks, &scal->sink, tile_ctr);
    if (iters == 0) GHET_CHECK(fill<double>(rank, inv_n, n, s));
    for (int32_t it = 0; it < iters; ++it) {
        const int last = (it == iters - 1);
        if (timing) GHET_HIP(hipEventRecord(kev[3 * it], s));
        if (L.blocked)
            hipLaunchKernelGGL(k_pr_hot, dim3(hot_blocks), dim3(kHotBT), 0, s, L.hot_words, L.hot_soff,
                               L.hot_tile_row, L.hot_tiles, L.hot_blocks, L.hot, L.outdeg, L.slot, long_part, ca, cb,
                               scal, static_cast<double>(n), damping, part, rank, last, tile_ctr);
        else if (hot_blocks > 0)
            hipLaunchKernelGGL(k_pr_iter, dim3(hot_blocks), dim3(kBT), 0, s, static_cast<const PRItem*>(L.items),
                               hot_items, L.rowptr, L.col, L.outdeg, L.slot, n, long_part, ca, cb, scal,
                               static_cast<double>(n), damping, part, rank, last);
        if (timing) GHET_HIP(hipEventRecord(kev[3 * it + 1], s));
        if (cold_blocks > 0)
            hipLaunchKernelGGL(k_pr_iter, dim3(cold_blocks), dim3(kBT), 0, s,
                               static_cast<const PRItem*>(L.items) + hot_items, cold_items, L.rowptr, L.col, L.outdeg,
                               L.slot, n, long_part, ca, cb, scal, static_cast<double>(n), damping, part + hot_parts,
                               rank, last);
        hipLaunchKernelGGL(k_pr_long, dim3(long_blocks), dim3(kBT), 0, s, static_cast<const int4*>(L.long_rows),
                           L.num_long_rows, long_part, L.outdeg, cb, scal, static_cast<double>(n), damping,
                           part + hot_parts + cold_blocks, rank, last);
        if (timing) GHET_HIP(hipEventRecord(kev[3 * it + 2], s));
        hipLaunchKernelGGL(k_pr_reduce, dim3(1), dim3(1024), 0, s, part, parts, &scal->sink, tile_ctr);
        std::swap(ca, cb);
    }
    const int norm = (normalise && iters > 0) ? 1 : 0;
    if (norm) {
        hipLaunchKernelGGL(k_pr_sum, dim3(init_blocks), dim3(kBT), 0, s, rank, n, part);
        hipLaunchKernelGGL(k_pr_reduce, dim3(1), dim3(1024), 0, s, part, init_blocks, &scal->total,
                           static_cast<int*>(nullptr));
    }
    hipLaunchKernelGGL(k_pr_unpermute, dim3(init_blocks), dim3(kBT), 0, s, rank, L.perm, n, scal, norm, out);
    GHET_HIP(hipEventRecord(g->ev_end, s));
    GHET_HIP(hipGetLastError());
    if (pr_out) GHET_CHECK(copy_out(g, pr_out, out, 8 * size_t(n), s));
    GHET_HIP(hipMemcpyAsync(g->pinned, scal, sizeof(PRScalars), hipMemcpyDeviceToHost, s));
    GHET_HIP(hipStreamSynchronize(s));
    if (st) {
        for (int32_t it = 0; it < iters; ++it) {
            float kms = 0, hms = 0;
            GHET_HIP(hipEventElapsedTime(&kms, kev[3 * it], kev[3 * it + 2]));
            GHET_HIP(hipEventElapsedTime(&hms, kev[3 * it], kev[3 * it + 1]));
            iter_ms += kms;
            hot_ms += hms;
        }
        float ms = 0;
        GHET_HIP(hipEventElapsedTime(&ms, g->ev_begin, g->ev_end));
        st->device_ms = ms;
        st->hot_kernel_ms = iter_ms;  // the iteration kernels (both phases and k_pr_long)
        st->hot_kernel_launches = iters;
        st->levels = iters;
        const PRScalars* hs = reinterpret_cast<const PRScalars*>(g->pinned);
        st->extra[0] = hs->sink;
        st->extra[1] = hs->total;
        st->extra[2] = phases ? hot_ms : 0;  // of which the hot phase (timed when the phases run apart)
        st->extra[3] = L.blocked ? 1 : 0;  // hot phase: 1 source-blocked, 0 per-edge gathers
    }
    return GHET_OK;
}

// Times k_gather_probe over the hot edges and over the cold edges of the
// PageRank layout (built as ghet_pr would build it for these options), each
// `reps` times; out4 = {median hot ms, median cold ms, hot edges, cold edges}.
int run_pr_gather_ceiling(ghet_graph* g, int32_t reps, const ghet_opts* o, double* out4) {
    if (!out4) return set_error(GHET_ERR_INVALID, "out4 is NULL");
    if (reps < 1 || reps > 1000) return set_error(GHET_ERR_INVALID, "reps = %d not in [1, 1000]", reps);
    if (!g->in.present() || g->h_in_rowptr.empty()) return set_error(GHET_ERR_INVALID, "PageRank needs the in-CSR");
    const int32_t n = g->n;
    for (int i = 0; i < 4; ++i) out4[i] = 0;
    if (n == 0) return GHET_OK;
    hipStream_t s = pick_stream(g, o);
    const int32_t split_code = (o && o->size >= sizeof(ghet_opts)) ? o->reserved[4] : 0;
    GHET_CHECK(build_pr_layout(g, s, split_code));
    const PRLayout& L = g->pr;
    int occ = 0, cus = 0;
    GHET_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(k_gather_probe), kBT, 0));
    GHET_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device));
    const int32_t grid = std::max(1, std::max(occ, 1) * std::max(cus, 1));
    void* p;
    float* contrib;
    GHET_CHECK(g->ws.get(WS_PR_CONTRIB, 4 * size_t(n), &p)); contrib = static_cast<float*>(p);
    GHET_CHECK(g->ws.get(WS_MISC, 4 * size_t(grid) * (kBT / 64), &p));
    float* out = static_cast<float*>(p);
    GHET_CHECK(fill<float>(contrib, 1.0f / static_cast<float>(n), n, s));
    const int64_t m = g->in.m, hot_e = L.num_hot_edges;
    hipEvent_t* kev = nullptr;
    GHET_CHECK(kernel_events(g, 2 * reps + 1, &kev));
    std::vector<double> ms[2];
    for (int phase = 0; phase < 2; ++phase) {
        const int64_t e0 = phase == 0 ? 0 : hot_e, e1 = phase == 0 ? hot_e : m;
        if (e1 <= e0) continue;
        GHET_HIP(hipEventRecord(kev[0], s));
        for (int32_t r = 0; r < reps; ++r) {
            hipLaunchKernelGGL(k_gather_probe, dim3(grid), dim3(kBT), 0, s, L.col, e0, e1, contrib, out);
            GHET_HIP(hipEventRecord(kev[1 + r], s));
        }
        GHET_HIP(hipGetLastError());
        GHET_HIP(hipStreamSynchronize(s));
        for (int32_t r = 0; r < reps; ++r) {
            float t = 0;
            GHET_HIP(hipEventElapsedTime(&t, kev[r], kev[1 + r]));
            ms[phase].push_back(t);
        }
    }
    for (int phase = 0; phase < 2; ++phase) {
        std::vector<double>& v = ms[phase];
        if (v.empty()) continue;
        std::sort(v.begin(), v.end());
        out4[phase] = v[v.size() / 2];
    }
    out4[2] = static_cast<double>(hot_e);
    out4[3] = static_cast<double>(m - hot_e);
    return GHET_OK;
}

}  // namespace ghet
