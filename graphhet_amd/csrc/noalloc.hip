// BFS that allocates nothing: for callers that cannot call hipMalloc (code
// under stream capture, tight serving loops, harnesses that forbid it).
//
// The caller passes a device out-CSR and a device depth array; the frontier
// queues and the level counters live in `scratch`, a file-scope device array
// of fixed size that each GPU holds from the first call until the process
// ends.  The traversal is level-synchronous top-down: every frontier vertex
// claims its unvisited out-neighbours with a CAS on their depth and appends
// the winners to the next queue.  Depths are unique, so the result equals
// ghet_bfs's whatever the schedule (0 at the source, INT32_MAX unreached;
// CPU_BFS.hpp:9-41).
#include <mutex>

#include "graph_internal.hpp"

namespace ghet {
namespace {

constexpr int32_t kNaMaxVertices = 1 << 22;  // queue capacity: a vertex enters a queue at most once
constexpr int kNaCounters = 64;              // level counters after the two queues
constexpr int kNaLight = 16;                 // rows up to this length are expanded lane-serially
constexpr int kNaMaxDevices = 64;

// Queue A [kNaMaxVertices], queue B [kNaMaxVertices], counters [kNaCounters]:
// 32 MiB + 256 B per GPU.
static __device__ __align__(256) int32_t scratch[2 * size_t(kNaMaxVertices) + kNaCounters];

__global__ void k_na_init(int32_t* __restrict__ depth, int32_t n, int32_t src, int32_t* __restrict__ queue,
                          int32_t* __restrict__ ctr) {
    for (int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < n;
         v += static_cast<int64_t>(gridDim.x) * blockDim.x)
        depth[v] = v == src ? 0 : kIntInf;
    if (blockIdx.x == 0) {
        if (threadIdx.x < kNaCounters) ctr[threadIdx.x] = 0;
        if (threadIdx.x == 0) queue[0] = src;
    }
}

__device__ __forceinline__ void na_visit(int32_t w, int32_t* depth, int32_t level, int32_t* next, int32_t* count) {
    const bool won = load_relaxed(depth + w) == kIntInf && atomicCAS(depth + w, kIntInf, level) == kIntInf;
    const int slot = wave_append(count, won);
    if (won) next[slot] = w;
}

// Expands the `cur` vertices of `queue` (level next_level - 1).  Each wave
// takes 64 entries; light rows run lane-serially, heavier rows with the whole
// wave.  Winners go to `next`, counted in `count`; `clear` (the counter the
// following level appends to) is zeroed here, after the host has read it.
__global__ void __launch_bounds__(256) k_na_expand(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                                   const int32_t* __restrict__ queue, int32_t cur, int32_t* depth,
                                                   int32_t next_level, int32_t* __restrict__ next, int32_t* count,
                                                   int32_t* clear) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *clear = 0;
    const int lane = lane_id();
    const int64_t groups = ceil_div(cur, 64);
    const int64_t wave = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const int64_t waves = (static_cast<int64_t>(gridDim.x) * blockDim.x) >> 6;
    for (int64_t g = wave; g < groups; g += waves) {
        const int64_t i = g * 64 + lane;
        int32_t b = 0, e = 0;
        if (i < cur) {
            const int32_t u = queue[i];
            b = rowptr[u];
            e = rowptr[u + 1];
        }
        if (e - b <= kNaLight)
            for (int32_t k = b; k < e; ++k) na_visit(col[k], depth, next_level, next, count);
        uint64_t heavy = __ballot(e - b > kNaLight);
        while (heavy) {
            const int srcl = __ffsll(static_cast<unsigned long long>(heavy)) - 1;
            heavy &= heavy - 1;
            const int32_t hb = __shfl(b, srcl), he = __shfl(e, srcl);
            for (int32_t k = hb + lane; k < he; k += 64) na_visit(col[k], depth, next_level, next, count);
        }
    }
}

std::mutex g_na_mu;                      // one call at a time per process: the scratch is shared
int32_t* g_na_scratch[kNaMaxDevices] = {};  // device address of `scratch` on each GPU

unsigned na_grid(int64_t threads) {
    return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(ceil_div(threads, 256), 8192)));
}

}  // namespace
}  // namespace ghet

using namespace ghet;

extern "C" int ghet_bfs_noalloc(int32_t n, int64_t m, const int32_t* d_rowptr, const int32_t* d_col, int32_t src,
                                int32_t* d_depth_out, void* stream) {
    clear_error();
    if (n < 0 || m < 0) return set_error(GHET_ERR_INVALID, "ghet_bfs_noalloc: n = %d, m = %lld", n, (long long)m);
    if (n > kNaMaxVertices)
        return set_error(GHET_ERR_UNSUPPORTED, "ghet_bfs_noalloc: n = %d exceeds its fixed scratch of %d vertices "
                         "(use ghet_bfs)", n, kNaMaxVertices);
    if (m > INT32_MAX) return set_error(GHET_ERR_UNSUPPORTED, "ghet_bfs_noalloc: m = %lld > 2^31 - 1", (long long)m);
    if (n == 0) return GHET_OK;
    if (src < 0 || src >= n) return set_error(GHET_ERR_INVALID, "ghet_bfs_noalloc: source %d not in [0, %d)", src, n);
    if (!d_rowptr || (m > 0 && !d_col) || !d_depth_out)
        return set_error(GHET_ERR_INVALID, "ghet_bfs_noalloc: NULL device array");
    int dev = 0;
    GHET_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNaMaxDevices) return set_error(GHET_ERR_UNSUPPORTED, "ghet_bfs_noalloc: device %d", dev);
    std::lock_guard<std::mutex> lock(g_na_mu);
    if (!g_na_scratch[dev]) {
        void* p = nullptr;
        GHET_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_na_scratch[dev] = static_cast<int32_t*>(p);
    }
    int32_t* const qa = g_na_scratch[dev];
    int32_t* const qb = qa + kNaMaxVertices;
    int32_t* const ctr = qb + kNaMaxVertices;
    hipStream_t s = static_cast<hipStream_t>(stream);

    hipLaunchKernelGGL(k_na_init, dim3(na_grid(n)), dim3(256), 0, s, d_depth_out, n, src, qa, ctr);
    GHET_HIP(hipGetLastError());
    int32_t cur = 1;
    for (int32_t level = 0; cur > 0; ++level) {
        int32_t* q = (level & 1) ? qb : qa;
        int32_t* nq = (level & 1) ? qa : qb;
        hipLaunchKernelGGL(k_na_expand, dim3(na_grid(int64_t(cur))), dim3(256), 0, s, d_rowptr, d_col, q, cur,
                           d_depth_out, level + 1, nq, ctr + (level & 1), ctr + ((level + 1) & 1));
        GHET_HIP(hipGetLastError());
        int32_t next = 0;
        GHET_HIP(hipMemcpyAsync(&next, ctr + (level & 1), sizeof(next), hipMemcpyDeviceToHost, s));
        GHET_HIP(hipStreamSynchronize(s));
        if (next < 0 || next > n)
            return set_error(GHET_ERR_HIP, "ghet_bfs_noalloc: level %d counted %d vertices (n = %d)", level + 1, next, n);
        cur = next;
    }
    return GHET_OK;
}
