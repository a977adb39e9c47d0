// Device graph handle: CSR upload into one 256-byte-aligned arena, borrowed
// device CSRs, the grow-only workspace and the C ABI dispatch.
//
// Replaces CSR_graph::allocate_device_memory / copy_host_to_device /
// release_device_memory (GPU_csr.hpp:120-215): instead of eight separate
// cudaMallocs, one arena holds every CSR array at 256-byte-aligned offsets,
// and the copies are queued on the graph's stream with one final sync.
#include <atomic>
#include <chrono>
#include <cstring>
#include <mutex>
#include <unordered_set>
#include <vector>

#include "graph_internal.hpp"

namespace ghet {
ghet_host_graph* host_graph_adopt(int32_t n, const int32_t* rowptr, const int32_t* col, const double* w,
                                  int threads);

namespace {
// Every graph and dynamic store handed out and not yet destroyed.  The set is
// allocated once and never freed, so a destroy that runs while the process
// exits still finds it.
struct LiveSet {
    std::mutex mu;
    std::unordered_set<void*> items[2];
};
LiveSet& live() {
    static LiveSet* s = new LiveSet;
    return *s;
}
}  // namespace

void live_add(int kind, void* p) {
    std::lock_guard<std::mutex> l(live().mu);
    live().items[kind].insert(p);
}
bool live_remove(int kind, void* p) {
    std::lock_guard<std::mutex> l(live().mu);
    return live().items[kind].erase(p) != 0;
}
std::vector<void*> live_take(int kind) {
    std::lock_guard<std::mutex> l(live().mu);
    return std::vector<void*>(live().items[kind].begin(), live().items[kind].end());
}

int Workspace::get(int slot, size_t want, void** out) {
    if (want == 0) want = 256;
    if (bytes[slot] < want) {
        const size_t rounded = (want + 4095) & ~size_t(4095);
        if (cap && static_cast<size_t>(total()) - bytes[slot] + rounded > cap)
            return set_error(GHET_ERR_NOMEM, "workspace slot %d needs %zu bytes: the graph's workspace cap of %zu bytes "
                             "would be exceeded (%lld held)", slot, rounded, cap, static_cast<long long>(total()));
        if (ptr[slot]) {
            hipFree(ptr[slot]);
            ptr[slot] = nullptr;
            bytes[slot] = 0;
        }
        size_t sz = (want + 4095) & ~size_t(4095);
        hipError_t e = hipMalloc(&ptr[slot], sz);
        if (e != hipSuccess) {
            ptr[slot] = nullptr;
            (void)hipGetLastError();
            return set_error(GHET_ERR_NOMEM, "workspace slot %d: hipMalloc(%zu) failed: %s", slot, sz,
                             hipGetErrorString(e));
        }
        bytes[slot] = sz;
    }
    *out = ptr[slot];
    return GHET_OK;
}

void Workspace::release() {
    for (int i = 0; i < WS_NUM_SLOTS; ++i) {
        if (ptr[i]) hipFree(ptr[i]);
        ptr[i] = nullptr;
        bytes[i] = 0;
    }
}

hipStream_t pick_stream(ghet_graph* g, const ghet_opts* opts) {
    if (opts && opts->size >= sizeof(ghet_opts) && opts->stream) return static_cast<hipStream_t>(opts->stream);
    return g->stream;
}

int kernel_events(ghet_graph* g, int count, hipEvent_t** out) {
    while (static_cast<int>(g->kevents.size()) < count) {
        hipEvent_t e;
        GHET_HIP(hipEventCreate(&e));
        g->kevents.push_back(e);
    }
    *out = g->kevents.data();
    return GHET_OK;
}

namespace {

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

int validate_csr(int32_t n, const int32_t* rowptr, const int32_t* col, const char* what) {
    if (!rowptr) return set_error(GHET_ERR_INVALID, "%s: rowptr is NULL", what);
    if (rowptr[0] != 0) return set_error(GHET_ERR_INVALID, "%s: rowptr[0] = %d, expected 0", what, rowptr[0]);
    for (int64_t v = 0; v < n; ++v)
        if (rowptr[v + 1] < rowptr[v])
            return set_error(GHET_ERR_INVALID, "%s: offsets not monotone at vertex %lld", what, (long long)v);
    const int64_t m = rowptr[n];
    if (m > 0 && !col) return set_error(GHET_ERR_INVALID, "%s: col is NULL", what);
    bool bad = false;
    int64_t bad_at = -1;
    std::mutex mu;
    parallel_for(m, hw_threads(0), [&](int64_t b, int64_t e, int) {
        for (int64_t i = b; i < e; ++i)
            if (col[i] < 0 || col[i] >= n) {
                std::lock_guard<std::mutex> l(mu);
                bad = true;
                bad_at = i;
                return;
            }
    });
    if (bad) return set_error(GHET_ERR_INVALID, "%s: column index out of range at edge %lld", what, (long long)bad_at);
    return GHET_OK;
}

int init_runtime(ghet_graph* g) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        (void)hipGetLastError();
        return set_error(GHET_ERR_NODEVICE, "no GPU visible to this process");
    }
    GHET_HIP(hipGetDevice(&g->device));
    GHET_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    GHET_HIP(hipEventCreate(&g->ev_begin));
    GHET_HIP(hipEventCreate(&g->ev_end));
    GHET_HIP(hipEventCreate(&g->ev_k0));
    GHET_HIP(hipEventCreate(&g->ev_k1));
    GHET_HIP(hipHostMalloc(reinterpret_cast<void**>(&g->pinned), kPinnedBytes, hipHostMallocDefault));
    return GHET_OK;
}

// Copies host arrays into device memory through a pinned staging buffer.
int upload(ghet_graph* g, void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return GHET_OK;
    GHET_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g->stream));
    return GHET_OK;
}

ghet_graph* new_graph() {
    auto* g = new ghet_graph();
    live_add(LIVE_GRAPH, g);
    return g;
}

// Frees a graph made by new_graph(); a pointer that is not live (already
// destroyed, or released by ghet_shutdown) is ignored.
void destroy_graph(ghet_graph* g) {
    if (!g || !live_remove(LIVE_GRAPH, g)) return;
    if (g->stream) hipStreamSynchronize(g->stream);
    g->ws.release();
    if (g->pr.mem) hipFree(g->pr.mem);
    if (g->pr.hot_mem) hipFree(g->pr.hot_mem);
    if (g->pr.items_mem) hipFree(g->pr.items_mem);
    if (g->prb.mem) hipFree(g->prb.mem);
    if (g->cdlp.mem) hipFree(g->cdlp.mem);
    if (g->cdlp.rmem) hipFree(g->cdlp.rmem);
    if (g->tc.mem) hipFree(g->tc.mem);
    if (g->tc.in_mem) hipFree(g->tc.in_mem);
    if (g->tc.task_mem) hipFree(g->tc.task_mem);
    if (g->ms.mem) hipFree(g->ms.mem);
    if (g->sl.mem) hipFree(g->sl.mem);
    if (g->bfs_skip) hipFree(g->bfs_skip);
    if (g->bfs_first) hipFree(g->bfs_first);
    if (g->arena) hipFree(g->arena);
    if (g->pinned) hipHostFree(g->pinned);
    if (g->mail) hipHostFree(g->mail);
    if (g->stage) hipHostFree(g->stage);
    if (g->ev_begin) hipEventDestroy(g->ev_begin);
    if (g->ev_end) hipEventDestroy(g->ev_end);
    if (g->ev_k0) hipEventDestroy(g->ev_k0);
    if (g->ev_k1) hipEventDestroy(g->ev_k1);
    for (hipEvent_t e : g->kevents) hipEventDestroy(e);
    if (g->stream) hipStreamDestroy(g->stream);
    delete g;
}

}  // namespace

int copy_out(ghet_graph* g, void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (bytes == 0) return GHET_OK;
    if (bytes < (size_t(4) << 20)) {
        GHET_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
        GHET_HIP(hipStreamSynchronize(s));
        return GHET_OK;
    }
    if (g->stage_bytes < bytes) {
        GHET_HIP(hipStreamSynchronize(s));  // the old buffer may still be a copy target
        if (g->stage) GHET_HIP(hipHostFree(g->stage));
        g->stage = nullptr;
        g->stage_bytes = 0;
        GHET_HIP(hipHostMalloc(&g->stage, bytes, hipHostMallocDefault));
        g->stage_bytes = bytes;
    }
    GHET_HIP(hipMemcpyAsync(g->stage, src, bytes, hipMemcpyDeviceToHost, s));
    GHET_HIP(hipStreamSynchronize(s));
    const char* from = static_cast<const char*>(g->stage);
    char* to = static_cast<char*>(dst);
    parallel_for(static_cast<int64_t>(bytes), std::min(hw_threads(0), 16),
                 [&](int64_t b, int64_t e, int) { std::memcpy(to + b, from + b, static_cast<size_t>(e - b)); });
    return GHET_OK;
}

int mail_init(ghet_graph* g) {
    if (g->mail) return GHET_OK;
    void* h = nullptr;
    GHET_HIP(hipHostMalloc(&h, sizeof(HostMail), hipHostMallocCoherent | hipHostMallocMapped));
    void* d = nullptr;
    const hipError_t e = hipHostGetDevicePointer(&d, h, 0);
    if (e != hipSuccess) {
        hipHostFree(h);
        GHET_HIP(e);
    }
    std::memset(h, 0, sizeof(HostMail));
    g->mail = static_cast<HostMail*>(h);
    g->mail_dev = static_cast<HostMail*>(d);
    g->mail_seq = 0;
    return GHET_OK;
}

int mail_wait(ghet_graph* g, hipStream_t s, unsigned long long seq) {
    const volatile unsigned long long* p = &g->mail->seq;
    const auto t0 = std::chrono::steady_clock::now();
    for (uint64_t k = 1;; ++k) {
        if (*p == seq) break;
        if ((k & 4095) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(20)) {
            GHET_HIP(hipStreamSynchronize(s));
            if (*p != seq) return set_error(GHET_ERR_HIP, "counter mailbox: sequence %llu not posted",
                                            static_cast<unsigned long long>(seq));
            break;
        }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    return GHET_OK;
}

// A directed graph owning an arena laid out as ghet_graph_create lays it out
// (out-CSR and weights, in-CSR and weights, all-CSR; absent parts have size
// 0 and null pointers).  The caller fills the device arrays and the host
// copies of the offsets.
int graph_alloc_arena(int32_t n, int64_t m_out, bool want_in, int64_t m_in, bool want_all, int64_t m_all,
                      bool want_w, ghet_graph** out) {
    *out = nullptr;
    auto* g = new_graph();
    g->n = n;
    g->directed = true;
    int st = init_runtime(g);
    if (st != GHET_OK) {
        destroy_graph(g);
        return st;
    }
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += align256(bytes); return o; };
    const size_t o_or = take(4 * (size_t(n) + 1)), o_oc = take(4 * size_t(m_out));
    const size_t o_ow = want_w ? take(8 * size_t(m_out)) : 0;
    size_t o_ir = 0, o_ic = 0, o_iw = 0, o_ar = 0, o_ac = 0;
    if (want_in) {
        o_ir = take(4 * (size_t(n) + 1));
        o_ic = take(4 * size_t(m_in));
        if (want_w) o_iw = take(8 * size_t(m_in));
    }
    if (want_all) {
        o_ar = take(4 * (size_t(n) + 1));
        o_ac = take(4 * size_t(m_all));
    }
    hipError_t e = hipMalloc(&g->arena, std::max<size_t>(off, 256));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        destroy_graph(g);
        return set_error(GHET_ERR_NOMEM, "device arena of %zu bytes: %s", off, hipGetErrorString(e));
    }
    g->arena_bytes = off;
    char* base = static_cast<char*>(g->arena);
    g->out = {n, m_out, reinterpret_cast<int32_t*>(base + o_or), reinterpret_cast<int32_t*>(base + o_oc),
              want_w ? reinterpret_cast<double*>(base + o_ow) : nullptr};
    if (want_in)
        g->in = {n, m_in, reinterpret_cast<int32_t*>(base + o_ir), reinterpret_cast<int32_t*>(base + o_ic),
                 want_w ? reinterpret_cast<double*>(base + o_iw) : nullptr};
    if (want_all)
        g->all = {n, m_all, reinterpret_cast<int32_t*>(base + o_ar), reinterpret_cast<int32_t*>(base + o_ac), nullptr};
    *out = g;
    return GHET_OK;
}

void graph_free(ghet_graph* g) { destroy_graph(g); }

}  // namespace ghet

using namespace ghet;

extern "C" {

int ghet_device_count(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return c;
}

int ghet_device_mem_info(size_t* free_bytes, size_t* total_bytes) {
    clear_error();
    if (!free_bytes || !total_bytes) return set_error(GHET_ERR_INVALID, "NULL output");
    *free_bytes = *total_bytes = 0;
    if (ghet_device_count() < 1) return set_error(GHET_ERR_NODEVICE, "no GPU visible");
    GHET_HIP(hipMemGetInfo(free_bytes, total_bytes));
    return GHET_OK;
}

int ghet_graph_create(int32_t n, const int32_t* out_rowptr, const int32_t* out_col, const double* out_w,
                      const int32_t* in_rowptr, const int32_t* in_col, const double* in_w,
                      const int32_t* all_rowptr, const int32_t* all_col, uint32_t flags, ghet_graph** out) {
    clear_error();
    if (!out) return set_error(GHET_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (n < 0) return set_error(GHET_ERR_INVALID, "n = %d < 0", n);
    GHET_CHECK(validate_csr(n, out_rowptr, out_col, "out-CSR"));
    const bool directed = !(flags & GHET_GRAPH_UNDIRECTED);
    const bool want_w = !(flags & GHET_GRAPH_NO_WEIGHTS);
    const bool want_in = directed && !(flags & GHET_GRAPH_NO_IN);
    const bool want_all = want_in && !(flags & GHET_GRAPH_NO_ALL);

    // Derive the in-/all-CSR on the host when the caller did not supply them.
    ghet_host_graph* derived = nullptr;
    ghet_csr_view vin{}, vall{};
    if (want_in && !in_rowptr) {
        derived = host_graph_adopt(n, out_rowptr, out_col, out_w, 0);
        if (!derived) return set_error(GHET_ERR_NOMEM, "host allocation failed");
        int st = ghet_host_graph_view(derived, GHET_VIEW_IN, &vin);
        if (st != GHET_OK) { ghet_host_graph_free(derived); return st; }
        in_rowptr = vin.rowptr;
        in_col = vin.col;
        in_w = vin.w;
    } else if (want_in) {
        int st = validate_csr(n, in_rowptr, in_col, "in-CSR");
        if (st != GHET_OK) return st;
    }
    if (want_all && !all_rowptr) {
        if (!derived) {
            derived = host_graph_adopt(n, out_rowptr, out_col, out_w, 0);
            if (!derived) return set_error(GHET_ERR_NOMEM, "host allocation failed");
        }
        int st = ghet_host_graph_view(derived, GHET_VIEW_ALL, &vall);
        if (st != GHET_OK) { ghet_host_graph_free(derived); return st; }
        all_rowptr = vall.rowptr;
        all_col = vall.col;
    } else if (want_all) {
        int st = validate_csr(n, all_rowptr, all_col, "all-CSR");
        if (st != GHET_OK) return st;
    }

    auto* g = new_graph();
    g->n = n;
    g->directed = directed;
    int st = init_runtime(g);
    if (st != GHET_OK) { destroy_graph(g); ghet_host_graph_free(derived); return st; }

    const int64_t m_out = out_rowptr[n];
    const int64_t m_in = want_in ? in_rowptr[n] : 0;
    const int64_t m_all = want_all ? all_rowptr[n] : 0;
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += align256(bytes); return o; };
    size_t o_or = take(4 * (size_t(n) + 1)), o_oc = take(4 * size_t(m_out));
    size_t o_ow = want_w ? take(8 * size_t(m_out)) : 0;
    size_t o_ir = 0, o_ic = 0, o_iw = 0, o_ar = 0, o_ac = 0;
    if (want_in) {
        o_ir = take(4 * (size_t(n) + 1));
        o_ic = take(4 * size_t(m_in));
        if (want_w) o_iw = take(8 * size_t(m_in));
    }
    if (want_all) {
        o_ar = take(4 * (size_t(n) + 1));
        o_ac = take(4 * size_t(m_all));
    }
    hipError_t e = hipMalloc(&g->arena, std::max<size_t>(off, 256));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        destroy_graph(g);
        ghet_host_graph_free(derived);
        return set_error(GHET_ERR_NOMEM, "device arena of %zu bytes: %s", off, hipGetErrorString(e));
    }
    g->arena_bytes = off;
    char* base = static_cast<char*>(g->arena);
    g->out = {n, m_out, reinterpret_cast<int32_t*>(base + o_or), reinterpret_cast<int32_t*>(base + o_oc),
              want_w ? reinterpret_cast<double*>(base + o_ow) : nullptr};
    g->h_out_rowptr.assign(out_rowptr, out_rowptr + n + 1);
    st = upload(g, g->out.rowptr, out_rowptr, 4 * (size_t(n) + 1));
    if (st == GHET_OK) st = upload(g, g->out.col, out_col, 4 * size_t(m_out));
    if (st == GHET_OK && want_w) {
        if (out_w) st = upload(g, g->out.w, out_w, 8 * size_t(m_out));
        else st = fill<double>(g->out.w, 1.0, m_out, g->stream);
    }
    if (want_in && st == GHET_OK) {
        g->in = {n, m_in, reinterpret_cast<int32_t*>(base + o_ir), reinterpret_cast<int32_t*>(base + o_ic),
                 want_w ? reinterpret_cast<double*>(base + o_iw) : nullptr};
        g->h_in_rowptr.assign(in_rowptr, in_rowptr + n + 1);
        st = upload(g, g->in.rowptr, in_rowptr, 4 * (size_t(n) + 1));
        if (st == GHET_OK) st = upload(g, g->in.col, in_col, 4 * size_t(m_in));
        if (st == GHET_OK && want_w) {
            if (in_w) st = upload(g, g->in.w, in_w, 8 * size_t(m_in));
            else st = fill<double>(g->in.w, 1.0, m_in, g->stream);
        }
    } else if (!directed) {
        g->in = g->out;
        g->h_in_rowptr = g->h_out_rowptr;
    }
    if (want_all && st == GHET_OK) {
        g->all = {n, m_all, reinterpret_cast<int32_t*>(base + o_ar), reinterpret_cast<int32_t*>(base + o_ac), nullptr};
        g->h_all_rowptr.assign(all_rowptr, all_rowptr + n + 1);
        st = upload(g, g->all.rowptr, all_rowptr, 4 * (size_t(n) + 1));
        if (st == GHET_OK) st = upload(g, g->all.col, all_col, 4 * size_t(m_all));
    } else if (!directed) {
        g->all = g->out;
        g->all.w = nullptr;
        g->h_all_rowptr = g->h_out_rowptr;
    }
    if (st == GHET_OK) {
        hipError_t se = hipStreamSynchronize(g->stream);
        if (se != hipSuccess) st = set_error(GHET_ERR_HIP, "upload: %s", hipGetErrorString(se));
    }
    ghet_host_graph_free(derived);
    if (st != GHET_OK) {
        destroy_graph(g);
        return st;
    }
    *out = g;
    return GHET_OK;
}

int ghet_graph_wrap_device(int32_t n, int64_t m_out, const int32_t* d_out_rowptr, const int32_t* d_out_col,
                           const double* d_out_w, int64_t m_in, const int32_t* d_in_rowptr, const int32_t* d_in_col,
                           int64_t m_all, const int32_t* d_all_rowptr, const int32_t* d_all_col, uint32_t flags,
                           ghet_graph** out) {
    clear_error();
    if (!out) return set_error(GHET_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (n < 0 || !d_out_rowptr || (m_out > 0 && !d_out_col))
        return set_error(GHET_ERR_INVALID, "wrap: out-CSR pointers missing");
    auto* g = new_graph();
    g->n = n;
    g->owns_csr = false;
    g->directed = !(flags & GHET_GRAPH_UNDIRECTED);
    int st = init_runtime(g);
    if (st != GHET_OK) { destroy_graph(g); return st; }
    g->out = {n, m_out, const_cast<int32_t*>(d_out_rowptr), const_cast<int32_t*>(d_out_col),
              const_cast<double*>(d_out_w)};
    if (d_in_rowptr) g->in = {n, m_in, const_cast<int32_t*>(d_in_rowptr), const_cast<int32_t*>(d_in_col), nullptr};
    else if (!g->directed) g->in = g->out;
    if (d_all_rowptr) g->all = {n, m_all, const_cast<int32_t*>(d_all_rowptr), const_cast<int32_t*>(d_all_col), nullptr};
    else if (!g->directed) g->all = g->out;
    // Host copies of the offsets feed the layout builders.
    auto pull = [&](std::vector<int32_t>& h, const int32_t* d) -> int {
        if (!d) return GHET_OK;
        h.resize(size_t(n) + 1);
        GHET_HIP(hipMemcpy(h.data(), d, 4 * (size_t(n) + 1), hipMemcpyDeviceToHost));
        return GHET_OK;
    };
    st = pull(g->h_out_rowptr, g->out.rowptr);
    if (st == GHET_OK) st = pull(g->h_in_rowptr, g->in.rowptr);
    if (st == GHET_OK) st = pull(g->h_all_rowptr, g->all.rowptr);
    // A negative edge count means "take it from the offsets".
    if (st == GHET_OK && m_out < 0) g->out.m = m_out = g->h_out_rowptr[n];
    if (st == GHET_OK && g->in.present() && g->directed && m_in < 0) g->in.m = g->h_in_rowptr[n];
    if (st == GHET_OK && g->all.present() && g->directed && m_all < 0) g->all.m = g->h_all_rowptr[n];
    if (st == GHET_OK && !g->directed) {
        g->in = g->out;
        g->all = g->out;
        g->all.w = nullptr;
    }
    if (st == GHET_OK && (g->h_out_rowptr[0] != 0 || g->h_out_rowptr[n] != m_out))
        st = set_error(GHET_ERR_INVALID, "wrap: out offsets do not span %lld edges", (long long)m_out);
    if (st != GHET_OK) { destroy_graph(g); return st; }
    *out = g;
    return GHET_OK;
}

void ghet_graph_destroy(ghet_graph* g) { destroy_graph(g); }

int ghet_shutdown(void) {
    clear_error();
    for (void* d : live_take(LIVE_DYN)) dyn_release(d);
    for (void* g : live_take(LIVE_GRAPH)) destroy_graph(static_cast<ghet_graph*>(g));
    mlp_release();
    return GHET_OK;
}

int ghet_graph_device_ptrs(const ghet_graph* g, ghet_device_ptrs* out) {
    clear_error();
    if (!g || !out) return set_error(GHET_ERR_INVALID, "NULL argument");
    out->n = g->n;
    out->m_out = g->out.m;
    out->m_in = g->in.m;
    out->m_all = g->all.m;
    out->out_rowptr = g->out.rowptr;
    out->out_col = g->out.col;
    out->out_w = g->out.w;
    out->in_rowptr = g->in.rowptr;
    out->in_col = g->in.col;
    out->in_w = g->in.w;
    out->all_rowptr = g->all.rowptr;
    out->all_col = g->all.col;
    return GHET_OK;
}

int ghet_graph_set_workspace_cap(ghet_graph* g, int64_t max_bytes) {
    clear_error();
    if (!g || max_bytes < 0) return set_error(GHET_ERR_INVALID, "graph handle is NULL or cap < 0");
    g->ws.cap = static_cast<size_t>(max_bytes);
    return GHET_OK;
}

int64_t ghet_graph_device_bytes(const ghet_graph* g) {
    if (!g) return 0;
    return static_cast<int64_t>(g->arena_bytes) + g->ws.total();
}

#define GHET_ENTER(g)                                                        \
    clear_error();                                                           \
    if (!(g)) return set_error(GHET_ERR_INVALID, "graph handle is NULL");    \
    if (stats) std::memset(stats, 0, sizeof(*stats));                        \
    if (hipSetDevice((g)->device) != hipSuccess) return set_error(GHET_ERR_HIP, "hipSetDevice failed")

int ghet_bfs(ghet_graph* g, int32_t src, int32_t* depth_out, const ghet_opts* opts, ghet_run_stats* stats) {
    GHET_ENTER(g);
    return run_bfs(g, src, depth_out, nullptr, opts, stats);
}
int ghet_bfs_pre(ghet_graph* g, int32_t src, int32_t* depth_out, int32_t* pre_out, const ghet_opts* opts,
                 ghet_run_stats* stats) {
    GHET_ENTER(g);
    if (!pre_out) return set_error(GHET_ERR_INVALID, "pre_out is NULL");
    return run_bfs(g, src, depth_out, pre_out, opts, stats);
}
int ghet_sssp(ghet_graph* g, int32_t src, double inf, double* dist_out, const ghet_opts* opts, ghet_run_stats* stats) {
    GHET_ENTER(g);
    return run_sssp(g, src, inf, dist_out, nullptr, opts, stats);
}
int ghet_sssp_pre(ghet_graph* g, int32_t src, double inf, double* dist_out, int32_t* pre_out, const ghet_opts* opts,
                  ghet_run_stats* stats) {
    GHET_ENTER(g);
    if (!pre_out) return set_error(GHET_ERR_INVALID, "pre_out is NULL");
    return run_sssp(g, src, inf, dist_out, pre_out, opts, stats);
}
int ghet_wcc(ghet_graph* g, int32_t* label_out, const ghet_opts* opts, ghet_run_stats* stats) {
    GHET_ENTER(g);
    return run_wcc(g, label_out, opts, stats);
}
int ghet_pr(ghet_graph* g, int32_t iterations, double damping, double* pr_out, const ghet_opts* opts,
            ghet_run_stats* stats) {
    GHET_ENTER(g);
    return run_pr(g, iterations, damping, pr_out, opts, stats);
}
int ghet_pr_gather_ceiling(ghet_graph* g, int32_t reps, const ghet_opts* opts, double* out4) {
    ghet_run_stats* stats = nullptr;
    GHET_ENTER(g);
    return run_pr_gather_ceiling(g, reps, opts, out4);
}
int ghet_cdlp(ghet_graph* g, int32_t iterations, int32_t* label_out, const ghet_opts* opts, ghet_run_stats* stats) {
    GHET_ENTER(g);
    return run_cdlp(g, iterations, label_out, opts, stats);
}
int ghet_tc(ghet_graph* g, int64_t* count_out, const ghet_opts* opts, ghet_run_stats* stats) {
    GHET_ENTER(g);
    return run_tc(g, count_out, opts, stats);
}

}  // extern "C"
