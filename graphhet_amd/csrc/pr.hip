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
// Layout (built once per graph from the in-offsets): consecutive rows are
// packed into tiles of at most kTileNnz in-edges and kTileRows rows; a row
// with more than kTileNnz in-edges is a "long row", split into kTileNnz-edge
// chunks whose partial sums are combined in chunk order by the last chunk to
// arrive.  One launch per iteration processes every tile and chunk:
//   1. stage contrib[in_col[e]] for the tile's contiguous edge range in LDS
//      (coalesced col reads, gathers that mostly hit L2/MALL);
//   2. rows with <= 32 staged values: one lane each; longer rows: one wave;
//   3. write contrib'[v] = (float)(pr'[v] / outdeg(v)) for the next
//      iteration, fold pr'[v] of sinks into a per-block FP64 partial, and on
//      the last iteration write pr'[v] in FP64.
// A one-block kernel then sums the partials in index order into S.
#include <cmath>

#include "graph_internal.hpp"

namespace ghet {
namespace {

constexpr int kBT = 256;
constexpr int kTileNnz = 2048;
constexpr int kTileRows = 1024;
constexpr int kStage = kTileNnz / kBT;  // staged values per thread
constexpr int kWaveRow = 32;            // rows with more staged values than this use a wave

struct PRScalars {
    double sink;   // S of the current ranks
    double total;  // sum of ranks (normalisation)
};

__device__ __forceinline__ double block_sum_f64(double x, double* s_red) {
    x = wave_sum(x);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) s_red[w] = x;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kBT / 64; ++i) t += s_red[i];
    __syncthreads();
    return t;  // valid in thread 0
}

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

__global__ void __launch_bounds__(kBT) k_pr_init(const int32_t* __restrict__ outdeg, int32_t n, double pr0,
                                                 float* __restrict__ contrib, double* __restrict__ part) {
    __shared__ double s_red[kBT / 64];
    double sink = 0;
    for (int64_t v = static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x; v < n;
         v += static_cast<int64_t>(gridDim.x) * kBT) {
        const int32_t d = outdeg[v];
        if (d > 0) contrib[v] = static_cast<float>(pr0 / d);
        else {
            contrib[v] = 0.f;
            sink += pr0;
        }
    }
    double t = block_sum_f64(sink, s_red);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// Sums part[0..count) in index order (strided lanes, then fixed tree) and
// stores it in *out.
__global__ void __launch_bounds__(1024) k_pr_reduce(const double* __restrict__ part, int32_t count, double* out) {
    __shared__ double s[1024];
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

__device__ __forceinline__ void finish_vertex(int32_t v, float sum, double base, double damping,
                                              const int32_t* __restrict__ outdeg, float* __restrict__ next,
                                              double* __restrict__ rank_out, bool last, double& sink) {
    const double pr = base + damping * static_cast<double>(sum);
    const int32_t d = outdeg[v];
    if (d > 0) next[v] = static_cast<float>(pr / d);
    else sink += pr;
    if (last) rank_out[v] = pr;
}

// One iteration over the relabeled in-CSR: blocks [0, num_tiles) are tiles,
// the rest are long-row chunks.
__global__ void __launch_bounds__(kBT) k_pr_iter(
    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col, const int32_t* __restrict__ outdeg,
    const int32_t* __restrict__ tile_beg, const int32_t* __restrict__ tile_end, int32_t num_tiles,
    const int32_t* __restrict__ long_row, const int32_t* __restrict__ long_chunk0, int32_t num_long_rows,
    int32_t* long_done, double* long_part, const float* __restrict__ contrib, float* __restrict__ next,
    const PRScalars* __restrict__ scal, double nd, double damping, double* __restrict__ part,
    double* __restrict__ rank_out, int last) {
    __shared__ float s_val[kTileNnz];
    __shared__ int32_t s_off[kTileRows + 1];
    __shared__ double s_red[kBT / 64];
    __shared__ float s_ks[kBT / 64], s_kc[kBT / 64];
    __shared__ int32_t s_last;
    const int tid = threadIdx.x;
    const double base = scal->sink * damping / nd + (1.0 - damping) / nd;
    double sink = 0;
    const int32_t b = blockIdx.x;
    if (b < num_tiles) {
        const int32_t rb = tile_beg[b], re = tile_end[b];
        const int32_t e0 = rowptr[rb], e1 = rowptr[re];
        const int32_t nnz = e1 - e0, rows = re - rb;
        // Stage: all column loads first, then all gathers (kStage in flight per lane).
        int32_t idx[kStage];
#pragma unroll
        for (int j = 0; j < kStage; ++j) {
            const int k = tid + j * kBT;
            idx[j] = k < nnz ? col[e0 + k] : -1;
        }
        float val[kStage];
#pragma unroll
        for (int j = 0; j < kStage; ++j) val[j] = idx[j] >= 0 ? contrib[idx[j]] : 0.f;
#pragma unroll
        for (int j = 0; j < kStage; ++j) {
            const int k = tid + j * kBT;
            if (k < nnz) s_val[k] = val[j];
        }
        for (int r = tid; r <= rows; r += kBT) s_off[r] = rowptr[rb + r] - e0;
        __syncthreads();
        for (int r = tid; r < rows; r += kBT) {
            const int32_t a = s_off[r], z = s_off[r + 1];
            if (z - a > kWaveRow) continue;
            Kahan32 acc;
            for (int32_t i = a; i < z; ++i) acc.add(s_val[i]);
            finish_vertex(rb + r, acc.value(), base, damping, outdeg, next, rank_out, last, sink);
        }
        // Rows with more staged values: wave w takes 64-row chunks w, w+4, ...
        // and reduces each such row cooperatively, in row order.
        const int wave = tid >> 6, lane = tid & 63;
        for (int c0 = wave * 64; c0 < rows; c0 += kBT) {
            const int r = c0 + lane;
            const bool wide = r < rows && (s_off[r + 1] - s_off[r]) > kWaveRow;
            uint64_t mask = __ballot(wide);
            while (mask) {
                const int rr = c0 + __ffsll(static_cast<unsigned long long>(mask)) - 1;
                mask &= mask - 1;
                const int32_t a = s_off[rr], z = s_off[rr + 1];
                Kahan32 acc;
                for (int32_t i = a + lane; i < z; i += 64) acc.add(s_val[i]);
                acc = wave_kahan(acc);
                if (lane == 0) finish_vertex(rb + rr, acc.value(), base, damping, outdeg, next, rank_out, last, sink);
            }
        }
        double t = block_sum_f64(sink, s_red);
        if (tid == 0) part[b] = t;
        return;
    }
    // Long-row chunk.
    const int32_t c = b - num_tiles;
    int lo = 0, hi = num_long_rows;  // largest L with long_chunk0[L] <= c
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (long_chunk0[mid] <= c) lo = mid;
        else hi = mid;
    }
    const int32_t L = lo, v = long_row[L];
    const int32_t k = c - long_chunk0[L], nchunks = long_chunk0[L + 1] - long_chunk0[L];
    const int32_t rs = rowptr[v], rz = rowptr[v + 1];
    const int32_t e0 = rs + k * kTileNnz, e1 = min(e0 + kTileNnz, rz);
    int32_t idx[kStage];
#pragma unroll
    for (int j = 0; j < kStage; ++j) {
        const int32_t i = e0 + tid + j * kBT;
        idx[j] = i < e1 ? col[i] : -1;
    }
    Kahan32 acc;
#pragma unroll
    for (int j = 0; j < kStage; ++j)
        if (idx[j] >= 0) acc.add(contrib[idx[j]]);
    acc = wave_kahan(acc);
    if ((tid & 63) == 0) {
        s_ks[tid >> 6] = acc.s;
        s_kc[tid >> 6] = acc.c;
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
        long_part[c] = static_cast<double>(t.s) + static_cast<double>(t.c);
        __threadfence();
        const int32_t arrived = atomicAdd(&long_done[L], 1);
        s_last = (arrived == nchunks - 1);
    }
    __syncthreads();
    if (tid == 0) {
        if (s_last) {
            __threadfence();
            double total = 0;
            const int32_t c0 = long_chunk0[L];
            for (int32_t j = 0; j < nchunks; ++j)
                total += __hip_atomic_load(&long_part[c0 + j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            long_done[L] = 0;
            const double pr = base + damping * total;
            const int32_t d = outdeg[v];
            if (d > 0) next[v] = static_cast<float>(pr / d);
            else sink += pr;
            if (last) rank_out[v] = pr;
        }
        part[b] = sink;
    }
}

__global__ void __launch_bounds__(kBT) k_pr_sum(const double* __restrict__ rank, int32_t n, double* __restrict__ part) {
    __shared__ double s_red[kBT / 64];
    double x = 0;
    for (int64_t v = static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x; v < n;
         v += static_cast<int64_t>(gridDim.x) * kBT)
        x += rank[v];
    double t = block_sum_f64(x, s_red);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// out[perm[r]] = rank[r] / total (total = 1 without normalisation).
__global__ void k_pr_unpermute(const double* __restrict__ rank, const int32_t* __restrict__ perm, int32_t n,
                               const PRScalars* scal, int normalise, double* __restrict__ out) {
    const double inv = normalise ? 1.0 / scal->total : 1.0;
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; r < n;
         r += static_cast<int64_t>(gridDim.x) * blockDim.x)
        out[perm[r]] = normalise ? rank[r] * inv : rank[r];
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

int build_pr_layout(ghet_graph* g, hipStream_t s) {
    PRLayout& L = g->pr;
    if (L.built) return GHET_OK;
    const int32_t n = g->n;
    const int64_t m = g->in.m;
    const std::vector<int32_t>& irp = g->h_in_rowptr;
    const std::vector<int32_t>& orp = g->h_out_rowptr;
    // Counting sort of vertices by out-degree, descending, stable by id.
    std::vector<int32_t> od(static_cast<size_t>(n)), perm(static_cast<size_t>(n)), inv(static_cast<size_t>(n));
    int32_t maxd = 0;
    for (int32_t v = 0; v < n; ++v) {
        od[v] = orp[v + 1] - orp[v];
        maxd = std::max(maxd, od[v]);
    }
    std::vector<int64_t> start(static_cast<size_t>(maxd) + 2, 0);
    for (int32_t v = 0; v < n; ++v) ++start[maxd - od[v] + 1];
    for (size_t i = 1; i < start.size(); ++i) start[i] += start[i - 1];
    for (int32_t v = 0; v < n; ++v) perm[start[maxd - od[v]]++] = v;
    for (int32_t r = 0; r < n; ++r) inv[perm[r]] = r;
    std::vector<int32_t> rp(static_cast<size_t>(n) + 1), odp(static_cast<size_t>(n));
    rp[0] = 0;
    for (int32_t r = 0; r < n; ++r) {
        const int32_t v = perm[r];
        rp[r + 1] = rp[r] + (irp[v + 1] - irp[v]);
        odp[r] = od[v];
    }
    // Tiles of consecutive relabeled rows; rows above kTileNnz become chunked long rows.
    std::vector<int32_t> tb, te, lrow, lch0;
    int32_t chunks = 0;
    for (int32_t r = 0; r < n;) {
        const int32_t d = rp[r + 1] - rp[r];
        if (d > kTileNnz) {
            lrow.push_back(r);
            lch0.push_back(chunks);
            chunks += static_cast<int32_t>(ceil_div(d, kTileNnz));
            ++r;
            continue;
        }
        const int32_t first = r;
        int32_t nnz = 0;
        while (r < n && r - first < kTileRows) {
            const int32_t dr = rp[r + 1] - rp[r];
            if (dr > kTileNnz || nnz + dr > kTileNnz) break;
            nnz += dr;
            ++r;
        }
        tb.push_back(first);
        te.push_back(r);
    }
    lch0.push_back(chunks);
    L.num_tiles = static_cast<int32_t>(tb.size());
    L.num_long_rows = static_cast<int32_t>(lrow.size());
    L.num_long_chunks = chunks;
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t bytes = al(4 * tb.size()) + al(4 * te.size()) + al(4 * lrow.size()) + al(4 * lch0.size()) +
                         al(4 * lrow.size()) + 3 * al(4 * size_t(n)) + al(4 * (size_t(n) + 1)) + al(4 * size_t(m)) + 256;
    GHET_HIP(hipMalloc(&L.mem, bytes));
    char* p = static_cast<char*>(L.mem);
    auto take = [&](size_t cnt) { int32_t* q = reinterpret_cast<int32_t*>(p); p += al(4 * cnt); return q; };
    L.tile_row = take(tb.size());
    L.tile_end_ptr = take(te.size());
    L.long_row = take(lrow.size());
    L.long_chunk0 = take(lch0.size());
    L.long_done = take(lrow.size());
    L.outdeg = take(size_t(n));
    L.perm = take(size_t(n));
    int32_t* d_inv = take(size_t(n));
    L.rowptr = take(size_t(n) + 1);
    L.col = take(size_t(m));
    auto up = [&](int32_t* dst, const std::vector<int32_t>& src) -> int {
        if (!src.empty()) GHET_HIP(hipMemcpyAsync(dst, src.data(), 4 * src.size(), hipMemcpyHostToDevice, s));
        return GHET_OK;
    };
    GHET_CHECK(up(L.tile_row, tb));
    GHET_CHECK(up(L.tile_end_ptr, te));
    GHET_CHECK(up(L.long_row, lrow));
    GHET_CHECK(up(L.long_chunk0, lch0));
    GHET_CHECK(up(L.outdeg, odp));
    GHET_CHECK(up(L.perm, perm));
    GHET_CHECK(up(d_inv, inv));
    GHET_CHECK(up(L.rowptr, rp));
    if (!lrow.empty()) GHET_CHECK(fill<int32_t>(L.long_done, 0, static_cast<int64_t>(lrow.size()), s));
    if (n > 0) {
        const unsigned grid = static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, kBT / 64), 65535)));
        hipLaunchKernelGGL(k_permute_rows, dim3(grid), dim3(kBT), 0, s, L.perm, d_inv, g->in.rowptr, g->in.col, L.rowptr,
                           L.col, n);
    }
    GHET_HIP(hipGetLastError());
    GHET_HIP(hipStreamSynchronize(s));
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
    GHET_CHECK(build_pr_layout(g, s));
    PRLayout& L = g->pr;
    const int32_t blocks = L.num_tiles + L.num_long_chunks;
    const int32_t init_blocks = static_cast<int32_t>(std::min<int64_t>(ceil_div(n, kBT), 4096));
    void* p;
    float *ca, *cb;
    double *rank, *out, *part, *long_part;
    PRScalars* scal;
    GHET_CHECK(g->ws.get(WS_PR_CONTRIB, 4 * size_t(n), &p)); ca = static_cast<float*>(p);
    GHET_CHECK(g->ws.get(WS_LABEL_A, 4 * size_t(n), &p)); cb = static_cast<float*>(p);
    GHET_CHECK(g->ws.get(WS_PR_RANK, 8 * size_t(n), &p)); rank = static_cast<double*>(p);
    GHET_CHECK(g->ws.get(WS_PR_OUT, 8 * size_t(n), &p)); out = static_cast<double*>(p);
    GHET_CHECK(g->ws.get(WS_PR_PART, 8 * size_t(std::max(blocks, init_blocks) + 1), &p)); part = static_cast<double*>(p);
    GHET_CHECK(g->ws.get(WS_PR_LONGPART, 8 * size_t(L.num_long_chunks + 1), &p)); long_part = static_cast<double*>(p);
    GHET_CHECK(g->ws.get(WS_COUNTERS, 256, &p)); scal = static_cast<PRScalars*>(p);
    const double inv_n = 1.0 / n;
    const bool timing = st != nullptr;
    double hot_ms = 0;
    hipEvent_t* kev = nullptr;
    if (timing) GHET_CHECK(kernel_events(g, 2 * iters, &kev));

    GHET_HIP(hipEventRecord(g->ev_begin, s));
    hipLaunchKernelGGL(k_pr_init, dim3(init_blocks), dim3(kBT), 0, s, L.outdeg, n, inv_n, ca, part);
    hipLaunchKernelGGL(k_pr_reduce, dim3(1), dim3(1024), 0, s, part, init_blocks, &scal->sink);
    if (iters == 0) GHET_CHECK(fill<double>(rank, inv_n, n, s));
    for (int32_t it = 0; it < iters; ++it) {
        const int last = (it == iters - 1);
        if (timing) GHET_HIP(hipEventRecord(kev[2 * it], s));
        hipLaunchKernelGGL(k_pr_iter, dim3(blocks), dim3(kBT), 0, s, L.rowptr, L.col, L.outdeg, L.tile_row,
                           L.tile_end_ptr, L.num_tiles, L.long_row, L.long_chunk0, L.num_long_rows, L.long_done,
                           long_part, ca, cb, scal, static_cast<double>(n), damping, part, rank, last);
        if (timing) GHET_HIP(hipEventRecord(kev[2 * it + 1], s));
        hipLaunchKernelGGL(k_pr_reduce, dim3(1), dim3(1024), 0, s, part, blocks, &scal->sink);
        std::swap(ca, cb);
    }
    const int norm = (normalise && iters > 0) ? 1 : 0;
    if (norm) {
        hipLaunchKernelGGL(k_pr_sum, dim3(init_blocks), dim3(kBT), 0, s, rank, n, part);
        hipLaunchKernelGGL(k_pr_reduce, dim3(1), dim3(1024), 0, s, part, init_blocks, &scal->total);
    }
    hipLaunchKernelGGL(k_pr_unpermute, dim3(init_blocks), dim3(kBT), 0, s, rank, L.perm, n, scal, norm, out);
    GHET_HIP(hipEventRecord(g->ev_end, s));
    GHET_HIP(hipGetLastError());
    if (pr_out) GHET_HIP(hipMemcpyAsync(pr_out, out, 8 * size_t(n), hipMemcpyDeviceToHost, s));
    GHET_HIP(hipMemcpyAsync(g->pinned, scal, sizeof(PRScalars), hipMemcpyDeviceToHost, s));
    GHET_HIP(hipStreamSynchronize(s));
    if (st) {
        for (int32_t it = 0; it < iters; ++it) {
            float kms = 0;
            GHET_HIP(hipEventElapsedTime(&kms, kev[2 * it], kev[2 * it + 1]));
            hot_ms += kms;
        }
        float ms = 0;
        GHET_HIP(hipEventElapsedTime(&ms, g->ev_begin, g->ev_end));
        st->device_ms = ms;
        st->hot_kernel_ms = hot_ms;
        st->hot_kernel_launches = iters;
        st->levels = iters;
        const PRScalars* hs = reinterpret_cast<const PRScalars*>(g->pinned);
        st->extra[0] = hs->sink;
        st->extra[1] = hs->total;
        st->extra[2] = L.num_tiles;
        st->extra[3] = L.num_long_chunks;
    }
    return GHET_OK;
}

}  // namespace ghet
