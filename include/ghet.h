/*
 * ghet.h -- C ABI of the MI355X (gfx950) CSR graph-analytics library.
 *
 * Plain C, POD arguments only: pointers, sizes and status codes.  No HIP or
 * torch types appear in any signature; a stream is passed as `void*`
 * (a hipStream_t, or NULL for the library's own stream).
 *
 * Every entry point returns a ghet_status.  On failure, ghet_last_error()
 * returns a message for the calling thread.  No entry point throws.
 *
 * Result buffers are caller-allocated HOST memory.  Passing NULL for a
 * result buffer runs the algorithm and leaves the result on the device only
 * (used for timing; ghet_run_stats still reports the device time).
 *
 * Semantics follow the reference's CPU adjacency-list algorithms; the
 * entry point each function replaces is cited beside it (paths relative to
 * the GraphHET source tree).
 */
#ifndef GHET_H
#define GHET_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GHET_ABI_VERSION 1

typedef enum ghet_status {
    GHET_OK = 0,
    GHET_ERR_INVALID = 1,     /* bad argument: NULL, source out of range, non-monotone offsets */
    GHET_ERR_NOMEM = 2,       /* host/device allocation failed or workspace cap exceeded */
    GHET_ERR_HIP = 3,         /* HIP runtime error (message in ghet_last_error) */
    GHET_ERR_NODEVICE = 4,    /* no GPU visible to this process */
    GHET_ERR_UNSUPPORTED = 5, /* input the path does not handle (e.g. > 2^31-1 edges) */
    GHET_ERR_IO = 6,          /* a file could not be opened, read or written, or is malformed */
    GHET_WARN_FALLBACK = 100  /* learned schedule unavailable: heuristic used, results valid */
} ghet_status;

/* Message of the last failing call on this thread ("" if none). */
const char* ghet_last_error(void);
int ghet_abi_version(void);
/* Number of visible GPUs (0 when none); never initialises more than HIP's runtime. */
int ghet_device_count(void);
/* Free and total device memory of the current GPU (replaces the LDBC driver's
 * cudaMemGetInfo calls, LDBC_GPU_csr.cu:74,116). */
int ghet_device_mem_info(size_t* free_bytes, size_t* total_bytes);

/* ------------------------------------------------------------------------ */
/* Host-side graphs (no GPU needed)                                          */
/* ------------------------------------------------------------------------ */

/* A read-only CSR view.  rowptr has n+1 entries; col/w have rowptr[n].
 * w may be NULL (unweighted). */
typedef struct ghet_csr_view {
    int32_t n;
    int64_t m;
    const int32_t* rowptr;
    const int32_t* col;
    const double* w;
} ghet_csr_view;

typedef struct ghet_host_graph ghet_host_graph;

enum ghet_view_kind {
    GHET_VIEW_OUT = 0, /* out-adjacency, neighbours ascending (GraphHET OUTs)            */
    GHET_VIEW_IN = 1,  /* in-adjacency (GraphHET INs), weights of the u->v edge          */
    GHET_VIEW_ALL = 2, /* per vertex: in-list then out-list (GraphHET all_Edges)         */
    GHET_VIEW_SYM = 3  /* undirected simple symmetrisation: union of in/out, no loops    */
};

/* Seeded R-MAT: ef*2^scale samples, each choosing `scale` quadrants with
 * probabilities (a,b,c,1-a-b-c).  Self-loops dropped, (u,v) duplicates merged
 * (the weight of the latest sample wins, as repeated add_edge does in
 * CPU_adj_list.hpp:110-122).  Weights (k+1)*2^-53, k uniform in [0,2^53).
 * threads<=0 picks the hardware thread count.  The graph is a function of
 * (scale, edge_factor, a, b, c, seed) only, never of `threads`. */
int ghet_gen_rmat(int scale, double edge_factor, double a, double b, double c,
                  uint64_t seed, int threads, ghet_host_graph** out);
/* Erdos-Renyi G(n, m): uniform (u,v) pairs, no loops, deduplicated, topped up
 * until exactly m distinct edges exist. */
int ghet_gen_er(int32_t n, int64_t m, uint64_t seed, int threads, ghet_host_graph** out);
/* Build from an edge list (duplicates: later entry's weight wins; self-loops kept). */
int ghet_host_graph_from_edges(int32_t n, int64_t m, const int32_t* src, const int32_t* dst,
                               const double* w, int threads, ghet_host_graph** out);
/* Build from a CSR (row offsets n+1, targets, optional weights): rows already
 * ascending without duplicates are adopted; others are canonicalised as
 * ghet_host_graph_from_edges does (the later entry's weight wins). */
int ghet_host_graph_from_csr(int32_t n, const int32_t* rowptr, const int32_t* col, const double* w,
                             int threads, ghet_host_graph** out);
/* Views are built on first request and owned by the graph. */
int ghet_host_graph_view(ghet_host_graph* g, int kind, ghet_csr_view* out);
void ghet_host_graph_free(ghet_host_graph* g);

/* Binary CSR cache: the out-adjacency in the vector-of-vectors layout of
 * binary_save_vector_of_vectors (binary_save_read_vector_of_vectors.hpp:10-36,
 * reproduced in include/CPU_adj_list/binary_save_read_vector_of_vectors.hpp):
 * an int32 row count, then per row an int32 length and that many elements.
 * `path` holds the targets (int32), `path` + ".w" the weights (double), so
 * each file reads back with binary_read_vector_of_vectors<int> / <double>.
 * Load checks both files against each other and the targets against the row
 * count.  Rows already ascending without duplicates are adopted as they are;
 * others are canonicalised as ghet_host_graph_from_edges does (sorted, the
 * later entry's weight wins).  A missing ".w" file gives unit weights.
 * I/O failures return GHET_ERR_IO with the file name in ghet_last_error. */
int ghet_host_graph_save(ghet_host_graph* g, const char* path);
int ghet_host_graph_load(const char* path, int threads, ghet_host_graph** out);

/* ------------------------------------------------------------------------ */
/* Device graphs                                                             */
/* ------------------------------------------------------------------------ */

typedef struct ghet_graph ghet_graph;

enum ghet_graph_flags {
    GHET_GRAPH_UNDIRECTED = 1u, /* in/all alias out (GPU_csr.hpp:133-140)           */
    GHET_GRAPH_NO_WEIGHTS = 2u, /* do not upload weights (SSSP then unavailable)     */
    GHET_GRAPH_NO_IN = 4u,      /* do not build/upload the in- and all-CSR           */
    GHET_GRAPH_NO_ALL = 8u      /* do not build/upload the all-CSR (CDLP unavailable) */
};

/* Device CSR handle: replaces toCSR's allocate/copy (GPU_csr.hpp:120-165,282-339).
 * out_* is required.  in_* and all_* may be NULL: they are then derived on the
 * host from out_*.  All buffers are copied; the caller keeps ownership of its
 * arrays.  Device buffers are 256-byte aligned sub-allocations of one arena. */
int ghet_graph_create(int32_t n,
                      const int32_t* out_rowptr, const int32_t* out_col, const double* out_w,
                      const int32_t* in_rowptr, const int32_t* in_col, const double* in_w,
                      const int32_t* all_rowptr, const int32_t* all_col,
                      uint32_t flags, ghet_graph** out);
/* Borrow existing device CSR arrays (the raw-core signatures of the reference
 * take device pointers: GPU_BFS_csr.cuh:258, GPU_SSSP_csr.cuh:253,
 * GPU_TC_csr.cuh:46).  Any pointer group except out may be NULL; a negative
 * edge count is read from the group's offsets.  Nothing is freed on destroy
 * except the library's own derived layouts and workspace. */
int ghet_graph_wrap_device(int32_t n,
                           int64_t m_out, const int32_t* d_out_rowptr, const int32_t* d_out_col,
                           const double* d_out_w,
                           int64_t m_in, const int32_t* d_in_rowptr, const int32_t* d_in_col,
                           int64_t m_all, const int32_t* d_all_rowptr, const int32_t* d_all_col,
                           uint32_t flags, ghet_graph** out);
/* Frees the graph.  A handle that is not live (destroyed before, or released
 * by ghet_shutdown) is ignored. */
void ghet_graph_destroy(ghet_graph* g);
/* Releases everything the library holds on the GPU: every live graph and
 * dynamic store (their streams, events, pinned and fine-grained host buffers,
 * layouts and workspaces) and the schedule networks' device memory.  Handles
 * given out before become invalid; destroying them afterwards is a no-op.
 * Call it before the process exits while the HIP runtime (and any profiler
 * attached to it) is still up; the Python binding registers it with atexit.
 * The library stays usable afterwards (schedule networks must be loaded
 * again).  No reference counterpart: the reference frees in destructors
 * (GPU_csr.hpp:167-215). */
int ghet_shutdown(void);

typedef struct ghet_device_ptrs {
    int32_t n;
    int64_t m_out, m_in, m_all;
    int32_t *out_rowptr, *out_col, *in_rowptr, *in_col, *all_rowptr, *all_col;
    double *out_w, *in_w;
} ghet_device_ptrs;
/* Device addresses of the CSR arrays (the CSR_graph public pointer fields). */
int ghet_graph_device_ptrs(const ghet_graph* g, ghet_device_ptrs* out);
/* Caps the graph's grow-only workspace (the per-call scratch slots, not the
 * CSR arena or the per-graph layouts) at max_bytes; 0 removes the cap.  A
 * call that would need more returns GHET_ERR_NOMEM with the sizes in
 * ghet_last_error and leaves the graph usable.  The reference has no cap: it
 * prints a failed cudaMalloc and continues (GPU_csr.hpp:21-25). */
int ghet_graph_set_workspace_cap(ghet_graph* g, int64_t max_bytes);
/* Device bytes held by the graph (CSR + layouts + workspace). */
int64_t ghet_graph_device_bytes(const ghet_graph* g);
/* On-GPU R-MAT: the graph ghet_gen_rmat(scale, edge_factor, a, b, c, seed)
 * followed by ghet_graph_create would give, bit for bit (same samples, loops
 * dropped, duplicates merged with the latest sample's weight, in-rows in
 * ascending source order, all-rows = in-list then out-list), generated in
 * device memory without a host copy.  flags: GHET_GRAPH_NO_WEIGHTS,
 * GHET_GRAPH_NO_IN, GHET_GRAPH_NO_ALL (UNDIRECTED is rejected).  No reference
 * counterpart: the reference generates and converts on the host
 * (CPU_adj_list.hpp:110-122, GPU_csr.hpp:282-339). */
int ghet_graph_gen_rmat(int scale, double edge_factor, double a, double b, double c, uint64_t seed,
                        uint32_t flags, ghet_graph** out);
/* Copies a device CSR (GHET_VIEW_OUT, _IN or _ALL) into caller buffers of
 * n+1 offsets and m entries; any pointer may be NULL; weights exist for OUT
 * and IN only.  Sizes: ghet_graph_device_ptrs. */
int ghet_graph_download(const ghet_graph* g, int kind, int32_t* rowptr, int32_t* col, double* w);


/* ------------------------------------------------------------------------ */
/* Options and statistics                                                    */
/* ------------------------------------------------------------------------ */

enum ghet_schedule {
    GHET_SCHED_HEURISTIC = 0, /* direction-optimising BFS, Delta heuristic for SSSP        */
    GHET_SCHED_PUSH = 1,      /* top-down queue expansion every level                      */
    GHET_SCHED_PULL = 2,      /* bottom-up every level after the first                     */
    GHET_SCHED_BITMAP = 3,    /* top-down over a frontier bitmap every level               */
    GHET_SCHED_REFERENCE = 4, /* reference rule: push if frontier < n/100, else pull       */
    GHET_SCHED_LEARNED = 5    /* MLP picks the mode (falls back to HEURISTIC if none)      */
};

enum ghet_bfs_mode { GHET_MODE_PUSH = 0, GHET_MODE_PULL = 1, GHET_MODE_BITMAP = 2 };

#define GHET_TRACE_FEATURES 16

/* One record per BFS level / SSSP phase when tracing is requested.
 * BFS: ms = device time of the level.  SSSP: one record per bucket-width
 * decision (before the first bucket and at every threshold advance); ms =
 * device time from the call's start to that decision, threshold = the
 * threshold at that moment (every vertex with a smaller distance is final). */
typedef struct ghet_level_trace {
    float features[GHET_TRACE_FEATURES]; /* inputs of the schedule network      */
    int32_t mode;                        /* mode or Delta class actually used   */
    float ms;                            /* see above                           */
    double threshold;                    /* SSSP: threshold at the decision     */
} ghet_level_trace;

typedef struct ghet_opts {
    uint32_t size;          /* sizeof(ghet_opts); 0 => defaults                   */
    int32_t schedule;       /* ghet_schedule                                      */
    double sssp_delta;      /* bucket width; <= 0 picks one                       */
    void* stream;           /* hipStream_t or NULL                                */
    ghet_level_trace* trace;/* optional per-level records                         */
    int32_t trace_cap;      /* capacity of `trace`                                */
    int32_t block_threads;  /* 0 = tuned default; else 64..1024 multiple of 64    */
    /* reserved[0] bit0: PR without the FP64 normalisation (reference form)
     *             bit1: PR runs its hot and cold gathers as two launches and
     *                   reports the hot phase's ms in stats extra[2]
     *             bit2: CDLP caps its hub merge tables at 64 slots (a test
     *                   setting: the merge then takes its sub-pass path)
     *             bit3: SSSP runs every heavy pass as a push over the
     *                   bucket's out-edges (no pull form)
     *             bit4: SSSP runs every heavy pass as a pull over the
     *                   in-edges of the unsettled vertices (a test setting;
     *                   needs the in-CSR); stats extra[3] counts the pulls
     * reserved[1] bit0: many-graph calls skip the autotuner
     * reserved[2]     : batched BFS pushes while frontier out-edges < m / this
     *                   (0 = default 16)
     * reserved[3]     : batched SSSP sources per pass (0 = 32, at most 32)
     * reserved[4]     : PR form and source windows: 0 = automatic;
     *                   3 << 16 = destination bins; 1 << 18 = pull gathers
     *                   with automatic windows; else pull gathers with
     *                   log2(hot sources) | log2(window) << 8, plus 1 << 16,
     *                   or 2 << 16 for the source-blocked hot phase (LDS
     *                   copies of the hot sources instead of per-edge
     *                   gathers); the layout is rebuilt when this changes
     * reserved[5]     : batched SSSP phase budget per pass (0 = automatic,
     *                   64 (n + 1024)); a pass that spends it, or two
     *                   minutes, stops and the call returns GHET_ERR_HIP
     * other entries must be 0                                             */
    int32_t reserved[6];
} ghet_opts;

typedef struct ghet_run_stats {
    double device_ms;          /* hipEvent time of the device work (no H2D/D2H)      */
    double hot_kernel_ms;      /* summed time of the path's dominant kernel          */
    int32_t hot_kernel_launches;
    int32_t levels;            /* BFS levels, SSSP phases, PR/CDLP iterations, ...   */
    int64_t edges_traversed;   /* m_r: out-edges of reached vertices (BFS/SSSP)      */
    int64_t vertices_reached;  /* n_r                                                */
    int32_t trace_len;         /* records written to opts->trace                     */
    int32_t status_flags;      /* bit0: learned fallback used                        */
    double extra[4];           /* algorithm-specific (PR: sink mass of last iter)    */
} ghet_run_stats;

/* ------------------------------------------------------------------------ */
/* Algorithms (single source / single graph)                                 */
/* ------------------------------------------------------------------------ */

/* BFS depths: 0 at src, INT32_MAX when unreached.
 * Replaces csr_bfs / Cuda_BFS (GPU_BFS_csr.cuh:258-335); semantics of
 * CPU_BFS (CPU_BFS.hpp:9-41). */
int ghet_bfs(ghet_graph* g, int32_t src, int32_t* depth_out,
             const ghet_opts* opts, ghet_run_stats* stats);
/* BFS with a deterministic parent: the smallest in-neighbour one level up
 * (-1 for src and unreached).  Replaces bfs_pre / Cuda_BFS_pre
 * (GPU_BFS_pre_csr.cuh:62-156). */
int ghet_bfs_pre(ghet_graph* g, int32_t src, int32_t* depth_out, int32_t* pre_out,
                 const ghet_opts* opts, ghet_run_stats* stats);
/* SSSP in FP64, unreached = inf (callers pass DBL_MAX).  Replaces csr_sssp /
 * Cuda_SSSP (GPU_SSSP_csr.cuh:253-327); semantics of CPU_shortest_paths
 * (CPU_shortest_paths.hpp:19-64).  Edge weights must be >= 0. */
int ghet_sssp(ghet_graph* g, int32_t src, double inf, double* dist_out,
              const ghet_opts* opts, ghet_run_stats* stats);
/* SSSP with a deterministic parent: the smallest u with fl(d[u]+w(u,v)) == d[v]
 * (-1 for src and unreached).  Replaces csr_sssp_pre / Cuda_SSSP_pre
 * (GPU_SSSP_pre_csr.cuh:256-368). */
int ghet_sssp_pre(ghet_graph* g, int32_t src, double inf, double* dist_out, int32_t* pre_out,
                  const ghet_opts* opts, ghet_run_stats* stats);
/* Weakly connected components; label = smallest vertex id in the component.
 * Replaces WCC_GPU / Cuda_WCC (GPU_WCC_csr.cuh:83-129). */
int ghet_wcc(ghet_graph* g, int32_t* label_out, const ghet_opts* opts, ghet_run_stats* stats);
/* PageRank, fixed iteration count, pr0 = 1/n, dangling mass redistributed.
 * Gather in FP32 with compensated per-vertex sums; dangling mass and the
 * stored rank in FP64.  Replaces GPU_PR / Cuda_PR (GPU_PR_csr.cuh:29-117,193-197);
 * semantics of PageRank (CPU_PageRank.hpp:12-56). */
int ghet_pr(ghet_graph* g, int32_t iterations, double damping, double* pr_out,
            const ghet_opts* opts, ghet_run_stats* stats);
/* The gather ceiling of ghet_pr's layout: times a kernel that only sums
 * contrib[col[e]] (FP32, eight loads in flight per thread, no rows) over the
 * hot edges and over the cold edges, `reps` times each.  out4 = {median hot
 * ms, median cold ms, hot edges, cold edges}.  No counterpart in the
 * reference: a measurement aid for callers comparing ghet_pr's phases (its
 * stats: extra[2] hot-phase ms) with the access pattern's own rate. */
int ghet_pr_gather_ceiling(ghet_graph* g, int32_t reps, const ghet_opts* opts, double* out4);
/* Synchronous label propagation over in+out neighbours with multiplicity,
 * ties to the smallest label, isolated vertices keep their label.
 * Replaces CDLP_GPU / Cuda_CDLP (GPU_CDLP_csr.cuh:75-210); semantics of
 * CDLP (CPU_Community_Detection.hpp:12-64). */
int ghet_cdlp(ghet_graph* g, int32_t iterations, int32_t* label_out,
              const ghet_opts* opts, ghet_run_stats* stats);
/* Per-vertex triangle counts in the reference's convention over the out-CSR:
 * for every u, v in N(u), w in N(u) and N(v): cnt[u]++, cnt[v]++, cnt[w]++
 * (CPU_TC.hpp:31-60).  On a symmetric loop-free graph this is 6*t(v).
 * Replaces Cuda_TC_PerNode / Cuda_TC (GPU_TC_csr.cuh:46-69). */
int ghet_tc(ghet_graph* g, int64_t* count_out, const ghet_opts* opts, ghet_run_stats* stats);

/* BFS depths that allocate nothing.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under stream capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The caller passes a
 * device out-CSR (d_rowptr: n+1 offsets, d_col: m targets) and a device
 * output d_depth_out[n]; depths are those of ghet_bfs (0 at src, INT32_MAX
 * unreached).  Frontier queues and counters live in a fixed device array of
 * the library (32 MiB + 256 B, held by each GPU the library runs on until the
 * process ends), so n is limited to 4,194,304 (2^22) vertices: larger graphs
 * return GHET_ERR_UNSUPPORTED.  Runs on `stream` (a hipStream_t, NULL for the
 * null stream) and returns when the depths are written.  Calls from one
 * process are serialised by a lock, since they share the array; the array
 * keeps the last call's contents, and each call clears what it needs.  The
 * device arrays are not validated (as with the reference's raw cores). */
int ghet_bfs_noalloc(int32_t n, int64_t m, const int32_t* d_rowptr, const int32_t* d_col, int32_t src,
                     int32_t* d_depth_out, void* stream);

/* ------------------------------------------------------------------------ */
/* Batched forms (no reference counterpart; results equal per-instance runs) */
/* ------------------------------------------------------------------------ */

/* k BFS sources in one persistent launch.  depth_out: k*n, row i = source i. */
int ghet_bfs_multi(ghet_graph* g, int32_t k, const int32_t* srcs, int32_t* depth_out,
                   const ghet_opts* opts, ghet_run_stats* stats);
/* k SSSP sources, up to 32 per pass, one persistent cooperative launch per
 * pass (the bucket loop runs on the device).  dist_out: k*n. */
int ghet_sssp_multi(ghet_graph* g, int32_t k, const int32_t* srcs, double inf, double* dist_out,
                    const ghet_opts* opts, ghet_run_stats* stats);

/* Many small graphs, packed: graph i owns vertices [voff[i], voff[i+1]) and
 * edges [rowptr[voff[i]], rowptr[voff[i+1]]) of one concatenated out-CSR whose
 * column ids are LOCAL to their graph.  WCC labels are local ids; TC counts use
 * the reference convention per graph.  Runs as one captured hipGraph. */
int ghet_wcc_many(int32_t count, const int32_t* voff, const int32_t* rowptr, const int32_t* col,
                  int32_t* label_out, const ghet_opts* opts, ghet_run_stats* stats);
int ghet_tc_many(int32_t count, const int32_t* voff, const int32_t* rowptr, const int32_t* col,
                 int64_t* count_out, const ghet_opts* opts, ghet_run_stats* stats);

/* ------------------------------------------------------------------------ */
/* Learned schedule and autotuner                                            */
/* ------------------------------------------------------------------------ */

/* Network 16 -> 64 -> 64 -> 4 (ReLU), FP16 weights, FP32 accumulate.
 * Parameter count: 16*64+64 + 64*64+64 + 64*4+4 = 5508. */
#define GHET_MLP_IN 16
#define GHET_MLP_H 64
#define GHET_MLP_OUT 4
#define GHET_MLP_PARAMS 5508

enum ghet_net_kind { GHET_NET_BFS = 0, GHET_NET_SSSP = 1 };

/* Load/save the schedule networks (a small text file).  A missing or corrupt
 * file leaves the heuristic in charge and returns GHET_WARN_FALLBACK. */
int ghet_sched_load(const char* path);
int ghet_sched_save(const char* path);
int ghet_sched_clear(void);
/* Install FP32 master parameters for one network (GHET_MLP_PARAMS floats). */
int ghet_sched_set_params(int kind, const float* params);
int ghet_sched_get_params(int kind, float* params);
/* Forward pass on the GPU: x [batch,16] FP32 -> logits [batch,4] FP32. */
int ghet_mlp_forward(int kind, int32_t batch, const float* x, float* logits);
/* The host mirror of the same forward pass (bit-identical logits); it makes
 * the per-level decisions inside ghet_bfs / ghet_sssp. */
int ghet_mlp_forward_host(int kind, int32_t batch, const float* x, float* logits);
/* One training step on the GPU (softmax cross-entropy, Adam).  Returns the
 * mean loss in *loss and, if grad_out != NULL, the FP32 gradient of the
 * loss w.r.t. the parameters before the update (GHET_MLP_PARAMS floats). */
int ghet_mlp_train_step(int kind, int32_t batch, const float* x, const int32_t* labels,
                        float lr, float* loss, float* grad_out);

/* Autotuner: times candidate block shapes for the batched launches and the
 * SSSP bucket width on graph g; results are cached in `cache_path` (may be NULL). */
int ghet_autotune(ghet_graph* g, const char* cache_path, int32_t* best_block_out);

/* ------------------------------------------------------------------------ */
/* Dynamic graphs                                                            */
/* ------------------------------------------------------------------------ */
/* The role of GraphHET's GPMA+ backend (GPU_gpma.hpp: class GPMA 42-73,
 * init_csr_gpma 928-941, init_gpma_from_csr 882-926, update_gpma 756-824):
 * a device edge store that takes batches of (src, dst, weight) updates with
 * GPMA's meaning -- weight 0 deletes the edge (absent: no-op), an existing
 * edge takes the new weight, any other edge is inserted -- and hands the
 * analytics a CSR.  Within one batch the last occurrence of an edge wins
 * (GPMA leaves repeated keys in one batch undefined).  Weight 0 cannot be
 * stored: ghet_dyn_create drops such edges and ghet_dyn_from_graph rejects
 * a graph that has them. */
typedef struct ghet_dyn ghet_dyn;
/* From a host CSR (rows canonicalised as ghet_host_graph_from_csr does). */
int ghet_dyn_create(int32_t n, const int32_t* rowptr, const int32_t* col, const double* w, int threads,
                    ghet_dyn** out);
/* From a device graph with its in-CSR (no host copy). */
int ghet_dyn_from_graph(const ghet_graph* g, ghet_dyn** out);
/* Applies one batch (w NULL: every weight 1.0).  stats: device_ms,
 * edges_traversed = distinct updates, levels = pool compactions, extra[0..3]
 * = inserted, deleted, modified, relocated slots. */
int ghet_dyn_update(ghet_dyn* d, int64_t count, const int32_t* src, const int32_t* dst, const double* w,
                    ghet_run_stats* stats);
int64_t ghet_dyn_edges(const ghet_dyn* d);
/* A ghet_graph holding the current edges (flags as ghet_graph_create:
 * NO_WEIGHTS, NO_IN, NO_ALL); every analytics entry point runs on it. */
int ghet_dyn_snapshot(ghet_dyn* d, uint32_t flags, ghet_graph** out);
void ghet_dyn_destroy(ghet_dyn* d);

#ifdef __cplusplus
}
#endif
#endif /* GHET_H */
