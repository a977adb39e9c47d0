// Weakly connected components on gfx950: Afforest-style sampled hooking,
// pointer jumping, and a full pass that skips the dominant component.
//
// Reference: WCC_GPU (GPU_WCC_csr.cuh:83-124) samples the first two in-
// neighbours, compresses, then links every remaining in-neighbour one thread
// per vertex; CPU_connected_components (CPU_connected_components.hpp:11-66)
// iterates label-min hooking to a fixpoint.  Both hook the higher root under
// the lower, so the final label of a vertex is the smallest vertex id of its
// component -- a closed form this kernel reproduces exactly: every hook here
// sets parent[h] = l with l < h and h a root at the time of the CAS, so
// parent chains strictly decrease and each tree's root is its minimum.
//
// The full pass reads the all-CSR (in-list ++ out-list), which holds every
// edge at both endpoints; that is what makes skipping the vertices already in
// the dominant component safe (an edge into it is seen from the other side).
#include "graph_internal.hpp"

namespace ghet {
namespace {

__device__ __forceinline__ int32_t find_root(const int32_t* parent, int32_t x) {
    for (;;) {
        int32_t p = load_relaxed(parent + x);
        if (p == x) return x;
        x = p;
    }
}

__device__ __forceinline__ void link(int32_t u, int32_t v, int32_t* parent) {
    int32_t a = find_root(parent, u), b = find_root(parent, v);
    while (a != b) {
        const int32_t h = a > b ? a : b, l = a > b ? b : a;
        const int32_t old = atomicCAS(parent + h, h, l);
        if (old == h) return;
        a = find_root(parent, old);
        b = find_root(parent, l);
    }
}

__global__ void k_wcc_init(int32_t* parent, int32_t n) {
    for (int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < n;
         v += static_cast<int64_t>(gridDim.x) * blockDim.x)
        parent[v] = static_cast<int32_t>(v);
}

// Link every vertex with its r-th neighbour in the given CSR.
__global__ void k_wcc_sample(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col, int32_t n,
                             int32_t r, int32_t* parent) {
    for (int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < n;
         v += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int32_t i = rowptr[v] + r;
        if (i < rowptr[v + 1]) link(static_cast<int32_t>(v), col[i], parent);
    }
}

__global__ void k_wcc_compress(int32_t* parent, int32_t n) {
    for (int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < n;
         v += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int32_t r = find_root(parent, static_cast<int32_t>(v));
        parent[v] = r;
    }
}

__global__ void k_wcc_gather(const int32_t* parent, const int32_t* idx, int32_t k, int32_t* out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < k) out[i] = parent[idx[i]];
}

// Link all neighbours from position r0 on; vertices whose root is `skip`
// are left out.  Light rows go lane-serial, heavier rows wave-cooperative.
__global__ void __launch_bounds__(256) k_wcc_full(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                                  int32_t n, int32_t r0, int32_t skip, int light_max, int32_t* parent) {
    const int lane = lane_id();
    const int64_t groups = ceil_div(n, 64);
    const int64_t wave_global = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const int64_t wave_stride = (static_cast<int64_t>(gridDim.x) * blockDim.x) >> 6;
    for (int64_t g = wave_global; g < groups; g += wave_stride) {
        const int64_t v = g * 64 + lane;
        int32_t b = 0, e = 0;
        bool act = false;
        if (v < n && load_relaxed(parent + v) != skip) {
            b = rowptr[v] + r0;
            e = rowptr[v + 1];
            act = b < e;
        }
        const bool light = act && (e - b) <= light_max;
        if (light)
            for (int32_t i = b; i < e; ++i) link(static_cast<int32_t>(v), col[i], parent);
        uint64_t heavy = __ballot(act && !light);
        while (heavy) {
            const int src = __ffsll(static_cast<unsigned long long>(heavy)) - 1;
            heavy &= heavy - 1;
            const int32_t hb = __shfl(b, src), he = __shfl(e, src);
            const int32_t u = static_cast<int32_t>(g * 64 + src);
            for (int32_t i = hb + lane; i < he; i += 64) link(u, col[i], parent);
        }
    }
}

unsigned grid_for(int64_t n, int per_block = 256) {
    return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, per_block), 16384)));
}

}  // namespace

int run_wcc(ghet_graph* g, int32_t* label_out, const ghet_opts* o, ghet_run_stats* st) {
    const int32_t n = g->n;
    if (!g->out.present()) return set_error(GHET_ERR_INVALID, "graph has no out-CSR");
    hipStream_t s = pick_stream(g, o);
    void* p;
    int32_t *parent, *samp;
    GHET_CHECK(g->ws.get(WS_PARENT, 4 * size_t(std::max(n, 1)), &p)); parent = static_cast<int32_t*>(p);
    constexpr int kSamples = 1024;
    GHET_CHECK(g->ws.get(WS_MISC, 8 * kSamples, &p)); samp = static_cast<int32_t*>(p);

    // Both directions of every edge: the all-CSR when present, else out + in,
    // else (undirected) the out-CSR alone.
    const bool use_all = g->all.present() && g->directed;
    const bool both = !use_all && g->directed && g->in.present();
    const bool can_skip = use_all || !g->directed || both;

    GHET_HIP(hipEventRecord(g->ev_begin, s));
    if (n > 0) {
        hipLaunchKernelGGL(k_wcc_init, dim3(grid_for(n)), dim3(256), 0, s, parent, n);
        const DevCSR& A = use_all ? g->all : g->out;
        for (int r = 0; r < 2; ++r) {
            hipLaunchKernelGGL(k_wcc_sample, dim3(grid_for(n)), dim3(256), 0, s, A.rowptr, A.col, n, r, parent);
            if (both)
                hipLaunchKernelGGL(k_wcc_sample, dim3(grid_for(n)), dim3(256), 0, s, g->in.rowptr, g->in.col, n, r,
                                   parent);
            hipLaunchKernelGGL(k_wcc_compress, dim3(grid_for(n)), dim3(256), 0, s, parent, n);
        }
        int32_t skip = -1;
        if (can_skip) {
            // The most frequent root among fixed, evenly spread sample vertices.
            std::vector<int32_t> idx(kSamples);
            for (int i = 0; i < kSamples; ++i)
                idx[i] = static_cast<int32_t>((static_cast<int64_t>(i) * 2654435761LL + 12345) % n);
            std::memcpy(g->pinned, idx.data(), sizeof(int32_t) * kSamples);
            GHET_HIP(hipMemcpyAsync(samp, g->pinned, sizeof(int32_t) * kSamples, hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(k_wcc_gather, dim3(kSamples / 256), dim3(256), 0, s, parent, samp, kSamples,
                               samp + kSamples);
            GHET_HIP(hipMemcpyAsync(g->pinned, samp + kSamples, sizeof(int32_t) * kSamples, hipMemcpyDeviceToHost, s));
            GHET_HIP(hipStreamSynchronize(s));
            std::vector<int32_t> roots(g->pinned, g->pinned + kSamples);
            std::sort(roots.begin(), roots.end());
            int best = 0;
            for (int i = 0, j; i < kSamples; i = j) {
                for (j = i; j < kSamples && roots[j] == roots[i]; ++j) {}
                if (j - i > best) {
                    best = j - i;
                    skip = roots[i];
                }
            }
        }
        const unsigned gw = grid_for(ceil_div(n, 64), 4);
        hipLaunchKernelGGL(k_wcc_full, dim3(gw), dim3(256), 0, s, A.rowptr, A.col, n, can_skip ? 2 : 0, skip, g->tune.wcc_light, parent);
        if (both)
            hipLaunchKernelGGL(k_wcc_full, dim3(gw), dim3(256), 0, s, g->in.rowptr, g->in.col, n, 2, skip, g->tune.wcc_light, parent);
        hipLaunchKernelGGL(k_wcc_compress, dim3(grid_for(n)), dim3(256), 0, s, parent, n);
    }
    GHET_HIP(hipEventRecord(g->ev_end, s));
    GHET_HIP(hipGetLastError());
    if (label_out && n > 0) GHET_CHECK(copy_out(g, label_out, parent, 4 * size_t(n), s));
    GHET_HIP(hipStreamSynchronize(s));
    if (st) {
        float ms = 0;
        GHET_HIP(hipEventElapsedTime(&ms, g->ev_begin, g->ev_end));
        st->device_ms = ms;
        st->levels = 1;
    }
    return GHET_OK;
}

}  // namespace ghet

namespace ghet {

// WCC without the sampled-mode host round trip (no component skipping): two
// sampled link rounds, compress, link everything, compress.  Nothing waits on
// the host, so the sequence can be captured into a hipGraph.
int enqueue_wcc_plain(ghet_graph* g, int32_t* parent, int light, int waves_per_cu, hipStream_t s) {
    const int32_t n = g->n;
    if (n == 0) return GHET_OK;
    const bool use_all = g->all.present() && g->directed;
    const bool both = !use_all && g->directed && g->in.present();
    const DevCSR& A = use_all ? g->all : g->out;
    hipLaunchKernelGGL(k_wcc_init, dim3(grid_for(n)), dim3(256), 0, s, parent, n);
    for (int r = 0; r < 2; ++r) {
        hipLaunchKernelGGL(k_wcc_sample, dim3(grid_for(n)), dim3(256), 0, s, A.rowptr, A.col, n, r, parent);
        if (both)
            hipLaunchKernelGGL(k_wcc_sample, dim3(grid_for(n)), dim3(256), 0, s, g->in.rowptr, g->in.col, n, r, parent);
        hipLaunchKernelGGL(k_wcc_compress, dim3(grid_for(n)), dim3(256), 0, s, parent, n);
    }
    const unsigned gw = static_cast<unsigned>(
        std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), int64_t(256) * waves_per_cu / 4)));
    hipLaunchKernelGGL(k_wcc_full, dim3(gw), dim3(256), 0, s, A.rowptr, A.col, n, 2, -1, light, parent);
    if (both) hipLaunchKernelGGL(k_wcc_full, dim3(gw), dim3(256), 0, s, g->in.rowptr, g->in.col, n, 2, -1, light, parent);
    hipLaunchKernelGGL(k_wcc_compress, dim3(grid_for(n)), dim3(256), 0, s, parent, n);
    GHET_HIP(hipGetLastError());
    return GHET_OK;
}

}  // namespace ghet
