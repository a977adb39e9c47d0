// Internal definition of the device graph handle and the per-algorithm
// launchers that the C ABI (abi.hip) dispatches to.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <memory>
#include <vector>

#include "dev_common.hpp"

namespace ghet {

constexpr size_t kPinnedBytes = 65536;

// A CSR resident in device memory.  rowptr has n+1 entries.
struct DevCSR {
    int32_t n = 0;
    int64_t m = 0;
    int32_t* rowptr = nullptr;
    int32_t* col = nullptr;
    double* w = nullptr;
    bool present() const { return rowptr != nullptr; }
};

// Grow-only device scratch, one buffer per slot, reused across calls on the
// same graph (the reference instead allocates and frees per call).
enum WsSlot {
    WS_DEPTH, WS_VISITED, WS_FRONT_A, WS_FRONT_B, WS_QUEUE_A, WS_QUEUE_B, WS_SCAN, WS_SCAN_AUX,
    WS_COUNTERS, WS_DIST, WS_STAMP_NEAR, WS_STAMP_FAR, WS_FAR_A, WS_FAR_B, WS_PARENT,
    WS_PR_CONTRIB, WS_PR_RANK, WS_PR_PART, WS_PR_LONGPART, WS_LABEL_A, WS_LABEL_B, WS_HASH,
    WS_TC_COUNT, WS_PRE, WS_PR_OUT, WS_MULTI_A, WS_MULTI_B, WS_MULTI_C, WS_MULTI_D, WS_MISC,
    WS_SSSP_LIST, WS_SSSP_START, WS_BFS_START, WS_SSSP_VERT, WS_BAR, WS_SSSP_TILE, WS_NUM_SLOTS
};

struct Workspace {
    void* ptr[WS_NUM_SLOTS] = {};
    size_t bytes[WS_NUM_SLOTS] = {};
    size_t cap = 0;  // 0: unlimited; else the most the slots may hold together
    // Returns a buffer of at least `want` bytes for `slot` (contents undefined).
    // GHET_ERR_NOMEM when the device allocation fails or the slots would
    // exceed `cap`; the slot keeps its old buffer in that case.
    int get(int slot, size_t want, void** out);
    void release();
    int64_t total() const {
        int64_t t = 0;
        for (size_t b : bytes) t += static_cast<int64_t>(b);
        return t;
    }
};

// Owns the temporary device buffers of one layout build.
struct DevTemp {
    std::vector<void*> ptrs;
    hipError_t err = hipSuccess;
    template <typename T>
    T* get(size_t count) {
        void* p = nullptr;
        if (err == hipSuccess) err = hipMalloc(&p, std::max<size_t>(count * sizeof(T), 256));
        if (err != hipSuccess) return nullptr;
        ptrs.push_back(p);
        return static_cast<T*>(p);
    }
    ~DevTemp() {
        if (!ptrs.empty()) (void)hipDeviceSynchronize();  // no launch may still read them
        for (void* p : ptrs) hipFree(p);
    }
};

// PageRank tile layout over the in-CSR (built once per graph).
struct PRLayout {
    bool built = false;
    int32_t hot = 0, window = 0;  // source split the layout was built with
    int32_t num_tiles = 0;        // tiles of consecutive short rows or segments
    int32_t num_long_chunks = 0;  // chunks of long rows and long segments
    int32_t num_long_rows = 0;    // rows finished from partials (k_pr_long)
    int32_t num_slots = 0;        // partial sums written per iteration
    int32_t num_segments = 0;     // cold (row, source window) segments
    int32_t num_items = 0;        // hot items, then one group per cold window
    int32_t num_hot_items = 0;    // items [0, num_hot_items) cover the hot parts
    int64_t num_hot_edges = 0;    // col[0, num_hot_edges) are the hot parts, the rest cold segments
    void* items = nullptr;        // [num_items] 32-byte work items (pr.hip), in items_mem
    void* items_mem = nullptr;
    void* long_rows = nullptr;    // [num_long_rows] int4 {row, first slot, slots, 0}
    int32_t* outdeg = nullptr;    // [n] out-degree of relabeled vertex r (sinks have 0)
    int32_t* perm = nullptr;      // [n] relabeled id -> caller's id (out-degree descending)
    int32_t* rowptr = nullptr;    // [n+S+1] offsets: n hot row parts, then S cold segments
    int32_t* slot = nullptr;      // [n+S] partial slot of a tile row / segment, -1: finish in place
    int32_t* col = nullptr;       // [m] relabeled sources: hot parts row-major, then cold segments
    void* mem = nullptr;
    // Source-blocked hot phase (blocked == true): the hot edges regrouped
    // into slices, one per (row tile, source block), each sorted by row; a
    // word packs (row in tile << 14 | source in block).
    bool blocked = false;
    int32_t hot_tiles = 0;        // row tiles
    int32_t hot_blocks = 0;       // source blocks
    uint32_t* hot_words = nullptr;  // [hot edges]
    int32_t* hot_soff = nullptr;    // [hot_tiles * hot_blocks + 1] slice offsets
    int32_t* hot_tile_row = nullptr;  // [hot_tiles + 1] first row of each tile
    void* hot_mem = nullptr;
};

// PageRank by destination bins (pr_bin.hip), built once per graph.  The
// edges are grouped by (source window of 2^15 ids, destination bin): window-
// major for the gather (a window's edges contiguous, one segment per bin
// inside it, each ordered by source) and bin-major for the accumulation (a
// bin's edges contiguous, its segments in window order).
struct PRBinLayout {
    bool built = false;
    bool unfit = false;        // the graph needs more bins than the tables hold (automatic form: gathers)
    int32_t num_win = 0;       // source windows
    int32_t num_bin = 0;       // destination bins (contiguous ids, <= 2^14 each)
    int32_t num_items = 0;     // gather items (an edge range of one window each)
    uint16_t* src = nullptr;   // [m] source - window start, window-major
    uint16_t* dst = nullptr;   // [m] destination - bin start, bin-major
    uint32_t* val = nullptr;   // [m + 1] one iteration's contribution words, bin-major (val[m]: spare)
    int2* segtab = nullptr;    // [num_win * num_bin + 1] {window-major start, bin-major start} of segment (w, b) at w * num_bin + b
    int32_t* binoff = nullptr;     // [num_bin + 1] first bin-major edge of each bin
    int32_t* bin_start = nullptr;  // [num_bin + 1] first destination of each bin
    int32_t* bin_order = nullptr;  // [num_bin] bins by edges, descending (launch order)
    int4* items = nullptr;     // [num_items] {window, first edge, end edge, first segment's bin}
    int32_t* item_nseg = nullptr;  // [num_items] segments the item's edges touch
    int32_t* outdeg = nullptr; // [n]
    void* mem = nullptr;
};

// CDLP runs its counting iterations in rank space: vertices ordered by
// all-degree, descending (ties by id), so the labels of the hubs, which most
// edges point at, sit together at the front of the label array.  rowptr_r /
// col_r are the all-CSR in rank space (each row's neighbours as ranks,
// ascending).  Label values stay vertex ids.  The degree classes (0 lane,
// 1 wave, 2 block with a 4K-slot LDS table, 3 block with a 16K-slot table,
// 4 hub, 5 no neighbours) are then contiguous rank ranges: list[c] is the
// range in rank space (a slice of iota), first[c] the same vertices by id,
// ascending, for the closed-form first iteration, which runs on the all-CSR
// itself.  A hub's edges are cut into 8192-edge chunks; each chunk
// writes its distinct (label, count) pairs to its own region of `pairs`,
// and merge items (hub, hash partition) add them up.
struct CDLPLayout {
    bool built = false;
    int32_t count[6] = {};
    int32_t* list[6] = {};
    int32_t* first[6] = {};
    int32_t* perm = nullptr;      // [n] rank -> vertex
    int32_t* iota = nullptr;      // [n] r -> r
    int32_t* rowptr_r = nullptr;  // [n + 1]
    int32_t* col_r = nullptr;     // [m_all]
    void* rmem = nullptr;         // perm, the first[] ids, iota, rowptr_r, col_r
    int32_t hub_chunks = 0;
    int32_t* chunk_hub = nullptr;    // [hub_chunks] index into list[4]
    int32_t* chunk_beg = nullptr;    // [hub_chunks] first edge of the chunk
    int32_t* hub_chunk0 = nullptr;   // [count[4] + 1] first chunk of each hub
    int32_t max_hub_chunks = 0;      // most chunks of one hub
    uint2* pairs = nullptr;          // [hub_chunks * 8192] {label, count}
    int32_t* chunk_pairs = nullptr;  // [hub_chunks] pairs a chunk wrote
    int32_t* hub_pairs = nullptr;    // [count[4]] pairs of each hub (zeroed by the plan)
    int32_t* item_off = nullptr;     // [count[4] + 1] first merge item of each hub; [count[4]] = items
    unsigned long long* hub_best = nullptr;  // [count[4]] packed (count, label) winner (zeroed by the final)
    void* mem = nullptr;
};

// Batched BFS: vertices whose in-lists are too long for one wave, and their
// in-edges cut into fixed-size chunks any wave can take.
// SSSP: the out-CSR with each row's edges ordered by weight (ties by target),
// so the edges with d[u] + w < thr form a prefix of the row.
struct SsspLayout {
    bool built = false;
    int32_t* col = nullptr;  // [m]
    double* w = nullptr;     // [m]
    void* mem = nullptr;
};

struct MSLayout {
    bool built = false;
    int32_t num_heavy = 0;
    int32_t num_chunks = 0;
    int32_t* heavy = nullptr;  // [num_heavy] vertex ids
    int2* chunks = nullptr;    // [num_chunks] {vertex, first in-edge}
    void* mem = nullptr;
};

// Degree-ordered DAG of a symmetric graph for TC.
struct TCLayout {
    bool built = false;
    bool symmetric = false;   // out-CSR symmetric and loop-free
    DevCSR dag;               // oriented by (degree, id); each row in that order
    int32_t* in_rowptr = nullptr;  // transpose of the counted CSR (DAG or out-CSR): in-lists
    int32_t* in_col = nullptr;
    int32_t* in_pos = nullptr;     // DAG: position of each in-entry's edge u -> v in u's row
    int* task_ctr = nullptr;       // k_tc_wave's task counter (zeroed before each count)
    void* in_mem = nullptr;
    void* wave_tasks = nullptr;    // {vertex, first in-list entry} per task (tc.hip)
    void* block_tasks = nullptr;
    int32_t num_wave_tasks = 0, num_block_tasks = 0;
    void* task_mem = nullptr;
    int64_t isect_bytes = 0;    // sum over DAG (or CSR) edges (u, v) of 4 (d+(u) + d+(v))
    void* mem = nullptr;
};

// Launch shapes the autotuner may change (results never depend on them).
struct Tuning {
    bool tuned = false;
    int tc_ept = 8;               // TC: rows per wave task
    int tc_waves_per_cu = 32;     // TC: grid size in waves per CU
    int wcc_light = 16;           // WCC: rows up to this length are linked lane-serially
    int wcc_waves_per_cu = 64;    // WCC: grid size in waves per CU
    int sssp_class = 1;           // SSSP: bucket width = base * 2^(class-1)
};

}  // namespace ghet

struct ghet_graph {
    int32_t n = 0;
    bool directed = true;
    bool owns_csr = true;
    int device = 0;
    ghet::DevCSR out, in, all;
    void* arena = nullptr;
    size_t arena_bytes = 0;
    std::vector<int32_t> h_out_rowptr, h_in_rowptr, h_all_rowptr;  // host copies of offsets
    ghet::Workspace ws;
    ghet::PRLayout pr;
    ghet::PRBinLayout prb;
    ghet::CDLPLayout cdlp;
    ghet::TCLayout tc;
    ghet::MSLayout ms;
    ghet::SsspLayout sl;
    double mean_w = -1.0;  // mean out-edge weight (SSSP bucket width), reduced on first use
    uint32_t* bfs_skip = nullptr;  // BFS: bit set for vertices no pull can reach (no in-edges), built on first use
    int32_t* bfs_first = nullptr;  // BFS: first in-neighbour of each vertex (-1 if none), built with bfs_skip
    ghet::Tuning tune;
    hipStream_t stream = nullptr;
    hipEvent_t ev_begin = nullptr, ev_end = nullptr, ev_k0 = nullptr, ev_k1 = nullptr;
    int32_t* pinned = nullptr;  // pinned host buffer (kPinnedBytes) for per-level counters
    void* stage = nullptr;              // pinned staging for result downloads (copy_out), grow-only
    size_t stage_bytes = 0;
    ghet::HostMail* mail = nullptr;      // fine-grained host mailbox (host address), made on first use
    ghet::HostMail* mail_dev = nullptr;  // its device address
    unsigned long long mail_seq = 0;     // last sequence number handed out
    std::vector<hipEvent_t> kevents;  // per-launch timing events, created on demand
};

namespace ghet {

// Registry of live handles, so that ghet_shutdown can release everything the
// library still holds before the HIP runtime is torn down at process exit.
enum LiveKind { LIVE_GRAPH = 0, LIVE_DYN = 1 };
void live_add(int kind, void* p);
bool live_remove(int kind, void* p);  // false if p was not live
std::vector<void*> live_take(int kind);
void dyn_release(void* d);  // dyn.hip: frees a live dynamic store
void mlp_release();         // mlp.hip: frees the schedule networks' device memory

// Copies `bytes` of device results to caller memory and returns when they
// are there.  Large results go through the graph's pinned staging buffer:
// one DMA at PCIe speed, then a multi-threaded host copy (a copy into
// pageable memory stages through small runtime buffers at about 6.5 GB/s).
int copy_out(ghet_graph* g, void* dst, const void* src, size_t bytes, hipStream_t s);
// Creates the graph's counter mailbox on first use.
int mail_init(ghet_graph* g);
// Next sequence number for a launch that posts to the mailbox.
inline unsigned long long mail_next(ghet_graph* g) { return ++g->mail_seq; }
// Waits until the mailbox holds sequence `seq` (polling; after 20 ms it
// synchronises the stream, which reports a failed launch, and checks once
// more).
int mail_wait(ghet_graph* g, hipStream_t s, unsigned long long seq);

// Resolves the stream to launch on (opts->stream or the graph's own stream).
hipStream_t pick_stream(ghet_graph* g, const ghet_opts* opts);
// At least `count` timing events owned by the graph.
int kernel_events(ghet_graph* g, int count, hipEvent_t** out);

// Launchers (return a ghet_status).  Output pointers are host memory or NULL.
int run_bfs(ghet_graph* g, int32_t src, int32_t* depth_out, int32_t* pre_out, const ghet_opts* o,
            ghet_run_stats* st);
int run_sssp(ghet_graph* g, int32_t src, double inf, double* dist_out, int32_t* pre_out, const ghet_opts* o,
             ghet_run_stats* st);
int run_wcc(ghet_graph* g, int32_t* label_out, const ghet_opts* o, ghet_run_stats* st);
int run_pr(ghet_graph* g, int32_t iters, double damping, double* pr_out, const ghet_opts* o, ghet_run_stats* st);
int run_pr_gather_ceiling(ghet_graph* g, int32_t reps, const ghet_opts* o, double* out4);
// PageRank by destination bins (pr_bin.hip); run_pr dispatches to it.
int run_pr_binned(ghet_graph* g, int32_t iters, double damping, double* pr_out, bool normalise, const ghet_opts* o,
                  ghet_run_stats* st);
// True when the binned form can hold this graph (source windows and bins fit its LDS tables).
bool pr_binned_fits(int32_t n, int64_t m);
// The automatic choice of form: destination bins for graphs past the L2-resident size.
bool pr_binned_default(int32_t n, int64_t m);
int run_cdlp(ghet_graph* g, int32_t iters, int32_t* label_out, const ghet_opts* o, ghet_run_stats* st);
int run_tc(ghet_graph* g, int64_t* count_out, const ghet_opts* o, ghet_run_stats* st);
int run_bfs_multi(ghet_graph* g, int32_t k, const int32_t* srcs, int32_t* depth_out, const ghet_opts* o,
                  ghet_run_stats* st);
// Capturable pieces (no host synchronisation inside) for hipGraph use.
int tc_prepare(ghet_graph* g, hipStream_t s);
int enqueue_tc(ghet_graph* g, unsigned long long* cnt, int ept, int waves_per_cu, hipStream_t s);
int enqueue_wcc_plain(ghet_graph* g, int32_t* parent, int light, int waves_per_cu, hipStream_t s);
int run_sssp_multi(ghet_graph* g, int32_t k, const int32_t* srcs, double inf, double* dist_out, const ghet_opts* o,
                   ghet_run_stats* st);

// Device-wide exclusive scan of int32 -> int32 (out may alias in); n up to 2^31-1.
// `aux` must hold scan_aux_ints(n) ints.  Writes the total to out[n].
int64_t scan_aux_ints(int64_t n);
int exclusive_scan_i32(const int32_t* in, int32_t* out, int64_t n, int32_t* aux, hipStream_t s);
int exclusive_scan_i64(const int64_t* in, int64_t* out, int64_t n, int64_t* aux, hipStream_t s);

// Reached-vertex statistics: n_r and m_r (sum of out-degrees) of vertices with
// depth != INT_MAX (or dist != inf), into slot fields 0 and 1 (host sums them).
int reached_stats_i32(const int32_t* depth, const int32_t* rowptr, int32_t n, CtrSlot* slots, hipStream_t s);
int reached_stats_f64(const double* dist, double inf, const int32_t* rowptr, int32_t n, CtrSlot* slots,
                      hipStream_t s);

template <typename T>
int fill(T* p, T v, int64_t n, hipStream_t s);

}  // namespace ghet
