// Triangle counting on gfx950: sorted-adjacency intersection, one wave per
// edge, binary search of the shorter list's elements in the longer list.
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
#include "graph_internal.hpp"

namespace ghet {
namespace {

constexpr int kBT = 256;
constexpr int kEdgesPerTask = 8;

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

// One wave per kEdgesPerTask consecutive edges (u, v) of X; for each, the
// common out-neighbours w of u and v: cnt[w] += 1, cnt[u] += c, cnt[v] += c.
__global__ void __launch_bounds__(kBT) k_tc_count(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                                  int32_t n, int64_t m, int ept, unsigned long long* cnt) {
    const int lane = lane_id();
    const int64_t tasks = ceil_div(m, ept);
    const int64_t wave_global = (static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x) >> 6;
    const int64_t wave_stride = (static_cast<int64_t>(gridDim.x) * kBT) >> 6;
    for (int64_t t = wave_global; t < tasks; t += wave_stride) {
        const int64_t e0 = t * ept, e1 = min(e0 + ept, m);
        int32_t u = upper_bound_i32(rowptr, static_cast<int64_t>(n) + 1, static_cast<int32_t>(e0)) - 1;
        int32_t ub = rowptr[u], ue = rowptr[u + 1];
        unsigned long long cu = 0;
        for (int64_t e = e0; e < e1; ++e) {
            while (ue <= e) {
                if (lane == 0 && cu) atomicAdd(&cnt[u], cu);
                cu = 0;
                ++u;
                ub = ue;
                ue = rowptr[u + 1];
            }
            const int32_t v = col[e];
            const int32_t vb = rowptr[v], ve = rowptr[v + 1];
            const bool u_small = (ue - ub) <= (ve - vb);
            const int32_t* small = col + (u_small ? ub : vb);
            const int32_t ns = u_small ? ue - ub : ve - vb;
            const int32_t* large = col + (u_small ? vb : ub);
            const int32_t nl = u_small ? ve - vb : ue - ub;
            unsigned long long c = 0;
            for (int32_t i = lane; i < ns; i += 64) {
                const int32_t w = small[i];
                if (contains(large, nl, w)) {
                    ++c;
                    atomicAdd(&cnt[w], 1ULL);
                }
            }
            c = wave_sum(c);
            if (lane == 0 && c) atomicAdd(&cnt[v], c);
            cu += c;
        }
        if (lane == 0 && cu) atomicAdd(&cnt[u], cu);
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
        GHET_HIP(hipStreamSynchronize(s));
        L.dag = {n, m_dag, drp, dcol, nullptr};
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
// so it can be captured into a hipGraph.  ept: edges per wave task;
// waves_per_cu: grid size in resident waves per CU.
int enqueue_tc(ghet_graph* g, unsigned long long* cnt, int ept, int waves_per_cu, hipStream_t s) {
    const int32_t n = g->n;
    const TCLayout& L = g->tc;
    const DevCSR& X = L.symmetric ? L.dag : g->out;
    GHET_CHECK(fill<uint64_t>(reinterpret_cast<uint64_t*>(cnt), 0ULL, n, s));
    if (X.m > 0) {
        const int64_t tasks = ceil_div(X.m, ept);
        const int64_t cap = int64_t(256) * waves_per_cu / (kBT / 64);
        hipLaunchKernelGGL(k_tc_count, dim3(grid_cap(std::min<int64_t>(ceil_div(tasks, kBT / 64), cap))), dim3(kBT), 0,
                           s, X.rowptr, X.col, n, X.m, ept, cnt);
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
