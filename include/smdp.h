/*
 * smdp.h -- C ABI of the MI355X (gfx950) neighbour-expansion library.
 *
 * The reference suite (NPSLab/show-me-dynamic-parallelism) exposes its hot path
 * as five zero-argument functions that talk through globals:
 *
 *   void SSSP_GPU();        apps/sssp/src/sssp.h:32
 *   void SpMV_GPU();        apps/SpMV/src/SpMV.h:38
 *   void PAGERANK_GPU();    apps/pagerank/src/pagerank.h:36
 *   void GRAPH_COLOR_GPU(); apps/graph-color/src/graph_color.h:24
 *   void BFS_REC_GPU();     apps/bfs-rec/src/bfs_rec.h:34
 *
 * Each one reads `graph`, `noNodeTotal`, `noEdgeTotal`, `source` and
 * `config.solution` (common/global.cpp:6-13, apps/sssp/src/sssp.h:13-22),
 * allocates device memory, runs, and copies its result back into a host
 * array.  Here the same calls take their inputs as arguments, return an error
 * code instead of calling exit(), and keep device buffers in a context that is
 * reused across calls.  No symbol takes or returns a torch type.
 *
 * Conventions
 *   - All indices are int32, as in the reference (common/global.h:28-42).
 *   - `smdp_csr` pointers are HOST pointers.  Results are copied into caller
 *     owned host arrays before the call returns (synchronous, like the
 *     reference wrappers).
 *   - The `_run` / `smdp_graph_*` family keeps the graph resident in HBM so a
 *     caller can time the device work alone; results stay on the device until
 *     smdp_graph_read_* copies them out.
 *   - Sentinels are the reference's: SSSP unreachable = 1073741824
 *     (common/global.h:19), BFS unreachable = 0x7FFFFFFF
 *     (apps/bfs-rec/src/bfs_rec.h:13), colour 0 = uncoloured and colours start
 *     at 1 (apps/graph-color/src/graph_color_wrapper.cu:118).
 *   - One context per host thread.  Calls on one context are serialised on
 *     the context's HIP stream.
 */
#ifndef SMDP_H
#define SMDP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes (the reference exit()s instead; see sssp_wrapper.cu:46-52,
 *      :250-253, :285-287 and sssp.cpp:114-122) ---------------------------- */
#define SMDP_OK              0
#define SMDP_ERR_ARG        -1  /* bad argument (NULL, negative size, bad source) */
#define SMDP_ERR_NOMEM      -2  /* device or host allocation failed */
#define SMDP_ERR_DEVICE     -3  /* no device, or a HIP runtime error */
#define SMDP_ERR_SOLUTION   -4  /* solution not supported by this app ("Unsupported solutions") */
#define SMDP_ERR_OVERFLOW   -5  /* a queue or round cap was exceeded ("Too many elements in queue") */
#define SMDP_ERR_IO         -6  /* file missing or malformed */
#define SMDP_ERR_FORMAT     -7  /* unknown input format ("Wrong code for dataset") */
#define SMDP_ERR_NOT_BUILT  -8  /* the library was built without device code */

/* ---- solutions: how high-degree ("heavy") vertices are handled ----------- */
/* The reference's -s switch (apps/sssp/src/sssp.cpp:31-36).  bfs-rec numbers
 * its switch differently (apps/bfs-rec/src/bfs_rec.cpp:30-36); the CLI maps it. */
#define SMDP_SOL_FLAT   0  /* one lane per vertex, whatever its degree (-s 0) */
#define SMDP_SOL_NAIVE  1  /* a heavy vertex is expanded at once by its whole wave (-s 1, per-thread child) */
#define SMDP_SOL_WAVE   2  /* heavy vertices go to a per-wave LDS queue drained by that wave (-s 2) */
#define SMDP_SOL_BLOCK  3  /* per-workgroup LDS queue drained by the workgroup after a barrier (-s 3) */
#define SMDP_SOL_GRID   4  /* grid-wide HBM queue drained by all workgroups in a second phase (-s 4) */
#define SMDP_SOL_HIER   5  /* BFS only: every frontier vertex expanded by a whole wave (bfs-rec -s 2) */
#define SMDP_SOL_TUNED  6  /* LDS-staged, edge-balanced tiles; rows/vertices above the split degree are
                              chunked over the grid; parameters from the autotuner */
#define SMDP_SOL_COUNT  7

/* ---- sentinels ----------------------------------------------------------- */
#define SMDP_SSSP_INF  1073741824      /* common/global.h:19 */
#define SMDP_BFS_UNDEF 0x7FFFFFFF      /* apps/bfs-rec/src/bfs_rec.h:13 */

typedef struct smdp_ctx smdp_ctx;
typedef struct smdp_graph smdp_graph;

/* A CSR graph in host memory.  row[n] == m.  w may be NULL when unused. */
typedef struct {
  int32_t n;
  int32_t m;
  const int32_t* row;
  const int32_t* col;
  const int32_t* w;
} smdp_csr;

/* Tunable launch parameters.  Zero in a field means "use the default". */
typedef struct {
  int32_t heavy_threshold;  /* degree at which a vertex is heavy (reference THRESHOLD = 64,
                               apps/sssp/src/sssp_kernel.cu:7) */
  int32_t split_threshold;  /* SMDP_SOL_TUNED: degree at which a vertex/row is chunked over the grid */
  int32_t chunk;            /* edges per grid chunk */
  int32_t wg_per_cu;        /* persistent workgroups per compute unit */
  int32_t batch_rounds;     /* rounds enqueued between two host checks of the queue length */
  int32_t max_rounds;       /* safety cap on rounds (SMDP_ERR_OVERFLOW when hit) */
  int32_t layout;           /* SpMV/PageRank tuned: arc layout, one of SMDP_LAYOUT_* (0 = automatic) */
  int32_t delta;            /* SSSP tuned, bucketed layout: bucket width, a power of two (0 = 64) */
  int32_t persistent;       /* SSSP tuned, bucketed layout: 0 = one launch per pass; 1 = every phase of the solve in
                               one launch whose workgroups are all resident, with a grid barrier between passes */
  int32_t group_share;      /* small-graph batch (read from the SMDP_APP_MSSSP parameters): a graph with more arcs
                               than group_share percent of one workgroup's fair share of the batch is solved by a
                               group of workgroups (0 = 100; 10..1000) */
  int32_t settle;           /* SSSP tuned, bucketed layout: how a bucket's settle pass reaches the arcs that leave
                               the bucket, one of SMDP_SETTLE_* (0 = automatic) */
} smdp_params;

/* smdp_params.settle.  Push: the bucket's vertices relax their out-arcs that
 * leave it.  Pull: every vertex beyond the bucket takes the smallest
 * candidate over its in-arcs from vertices below the bucket's end (needs an
 * in-arc copy of the graph, 8 B per arc, built once per resident graph).
 * Automatic: pull when those in-arcs are fewer than the bucket's out-arcs,
 * from degree counts gathered during the bucket's light passes. */
#define SMDP_SETTLE_AUTO 0
#define SMDP_SETTLE_PUSH 1
#define SMDP_SETTLE_PULL 2

/* Arc layouts of the tuned SpMV/PageRank solution (smdp_params.layout). */
#define SMDP_LAYOUT_AUTO    0  /* PageRank: BINNED when the graph has at least 2^20 arcs, else CSR; SpMV: CSR */
#define SMDP_LAYOUT_CSR     1  /* gather from the in-CSR as given */
#define SMDP_LAYOUT_BINNED  2  /* PageRank only: arcs regrouped once per graph by (source chunk, row bin); each
                                  iteration is two streaming passes through LDS instead of random gathers */
#define SMDP_LAYOUT_HOT     3  /* SSSP/BFS tuned: the solve runs on a copy of the graph with vertices renumbered by
                                  falling in-degree (built once per resident graph; results are reported in the
                                  caller's numbering).  Automatic picks it for resident graphs of >= 2^20 arcs. */
#define SMDP_LAYOUT_BUCKETS 4  /* SSSP tuned: bucketed rounds (delta-stepping) on the renumbered copy.  Vertices are
                                  taken in order of distance buckets of width delta; within a bucket, arcs that stay in
                                  the bucket are relaxed until no distance in it changes, then every other arc of the
                                  bucket's vertices is relaxed once.  Automatic picks it for resident graphs of
                                  >= 2^20 arcs without negative weights; a negative weight falls back to
                                  SMDP_LAYOUT_HOT. */
#define SMDP_LAYOUT_XCD8    8  /* sources split into 8 blocks, one per XCD, so each XCD gathers from a slice
                                  that fits its L2 */

/* Work counters and timings of the last call. */
typedef struct {
  double   ms_kernel;          /* device time of the compute region (HIP events on the context stream) */
  double   ms_total;           /* host wall time of the whole call, copies included */
  uint64_t edges_relaxed;      /* arcs scanned by the kernels (re-scans included) */
  uint64_t vertices_expanded;  /* frontier entries / rows processed */
  uint64_t rounds;             /* rounds (SSSP/BFS), iterations (PageRank), colours (coloring) */
  uint64_t q_push[4];          /* pushes into [0] next frontier, [1] wave queues, [2] workgroup
                                  queues, [3] grid queue */
  uint64_t m_reached;          /* out-arcs of reached vertices (the GTEPS numerator) */
  double   residual;           /* PageRank: last FP64 L1 residual */
  double   dangling;           /* PageRank: last dangling mass */
  uint64_t kernel_launches;    /* kernels launched by the call */
  double   ms_dominant;        /* summed device time of the phase-1 kernel (when profiling is on) */
  uint64_t dominant_launches;  /* launches of the phase-1 kernel (when profiling is on) */
  double   ms_phase2;          /* summed device time of the phase-2 (grid queue) kernel (profiling) */
  uint64_t phase2_launches;    /* launches of the phase-2 kernel (profiling) */
  uint64_t edges_phase2;       /* arcs of edges_relaxed scanned by the phase-2 kernel */
  uint64_t atomics_tried;      /* SSSP/BFS: atomicMin issued after a plain read said "improves" */
  uint64_t atomics_won;        /* SSSP/BFS: atomicMin that lowered the distance */
  uint64_t arcs_skipped;       /* bucketed SSSP: arcs read but left to the bucket's other pass (not relaxed,
                                  not in edges_relaxed) */
} smdp_stats;

/* PageRank options (the reference hard-codes them: DAMPING_FACTOR 0.5 and
 * EPSILON 1e-13 in apps/pagerank/src/pagerank_wrapper.cu:7-8, TOTAL_RANK 1000
 * in pagerank.h:15). */
typedef struct {
  float   damping;        /* d */
  float   total_rank;     /* T: ranks sum to T */
  double  tol_l1;         /* stop when sum|r'-r| / T < tol_l1 (FP64, on the device) */
  int32_t max_iter;       /* stop after this many iterations */
  int32_t init_zero;      /* 1: start from r = 0 as prepare_gpu does (pagerank_wrapper.cu:115);
                             0: start from r = T/N (pagerank.cpp:103-106) */
} smdp_pr_opts;

/* ---- context --------------------------------------------------------------*/
/* Replaces prepare_gpu()'s cudaSetDevice(config.device_num) (sssp_wrapper.cu:63). */
int         smdp_create(int device, smdp_ctx** out);
void        smdp_destroy(smdp_ctx* ctx);
const char* smdp_strerror(int code);
/* Text of the last HIP failure seen on this host thread ("" if none). */
const char* smdp_last_hip_error(void);
int         smdp_get_params(smdp_ctx* ctx, int app, smdp_params* out);
int         smdp_set_params(smdp_ctx* ctx, int app, const smdp_params* p);
/* name: device name; cus: compute units; arch: e.g. "gfx950". */
int         smdp_device_info(smdp_ctx* ctx, char* name, int name_len, char* arch, int arch_len,
                             int32_t* cus);
/* 1: record per-launch HIP events for the dominant kernel (fills ms_dominant)
 * and, for SSSP/BFS, a per-round log of queue lengths and arcs scanned. */
int         smdp_set_profiling(smdp_ctx* ctx, int on);
/* Per-round log of the last SSSP/BFS call made with profiling on: queue
 * length, arcs scanned and device time (phase 1 + phase 2) of round i.
 * Returns the number of rounds logged (at most cap), or a negative error. */
int         smdp_round_log(smdp_ctx* ctx, int64_t* qlen, int64_t* arcs, double* ms, int32_t cap);

/* apps, for smdp_get/set_params and the autotuner */
#define SMDP_APP_SSSP     0
#define SMDP_APP_BFS      1
#define SMDP_APP_SPMV     2
#define SMDP_APP_PAGERANK 3
#define SMDP_APP_COLOR    4
#define SMDP_APP_MSSSP    5
#define SMDP_APP_COUNT    6

/* ---- drop-in calls: host arrays in, host arrays out ------------------------*/
/* SSSP_GPU (apps/sssp/src/sssp_wrapper.cu:304-341): dist[v] for every v. */
int smdp_sssp(smdp_ctx* ctx, const smdp_csr* g, int32_t src, int solution,
              int32_t* dist, smdp_stats* st);
/* BFS_REC_GPU (apps/bfs-rec/src/bfs_rec_wrapper.cu:240-280): level[v]. */
int smdp_bfs(smdp_ctx* ctx, const smdp_csr* g, int32_t src, int solution,
             int32_t* level, smdp_stats* st);
/* SpMV_GPU (apps/SpMV/src/SpMV_wrapper.cu:193-224): y = A x, A in CSR with
 * values val[m]; FP32 in and out with compensated accumulation. */
int smdp_spmv_f32(smdp_ctx* ctx, const smdp_csr* a, const float* val, const float* x,
                  int solution, float* y, smdp_stats* st);
/* PAGERANK_GPU (apps/pagerank/src/pagerank_wrapper.cu:366-400): pull over
 * the in-CSR (childVertexArray / rEdgeArray, common/global.cpp:237-277);
 * outdeg[n] is the out-degree (pagerank.cpp:108-113). */
int smdp_pagerank_f32(smdp_ctx* ctx, const smdp_csr* in_csr, const int32_t* outdeg,
                      const smdp_pr_opts* opt, int solution, float* rank, smdp_stats* st);
/* GRAPH_COLOR_GPU (apps/graph-color/src/graph_color_wrapper.cu:210-238).
 * prio == NULL reproduces the reference's vertex-ID priority. */
int smdp_color(smdp_ctx* ctx, const smdp_csr* g, const uint32_t* prio, int solution,
               int32_t* color, smdp_stats* st);

/* Tree Height and Tree Descendant, the recursive benchmarks that the
 * reference's README.md:12-13 lists but does not ship.  g holds the child
 * lists of a tree (w is ignored); a vertex with two or more in-arcs anywhere
 * in g is SMDP_ERR_ARG, and arcs into the root are ignored.  For every
 * vertex reached from root: height[v] = 0 for a leaf, else 1 + the largest
 * child height; desc[v] = the number of its descendants.  Both are -1 for
 * vertices not reached.  height or desc may be NULL; *tree_height (may be
 * NULL) = height[root]. */
int smdp_tree(smdp_ctx* ctx, const smdp_csr* g, int32_t root, int32_t* height, int32_t* desc,
              int32_t* tree_height, smdp_stats* st);

/* ---- batched SSSP (no counterpart in the reference) ------------------------*/
/* k sources on one graph in one launch sequence; dist is [k][n] row-major. */
int smdp_sssp_multi_source(smdp_ctx* ctx, const smdp_csr* g, const int32_t* srcs, int32_t k,
                           int32_t* dist, smdp_stats* st);
/* ng independent graphs, one source each, solved in one captured hipGraph
 * launch.  dist[i] receives graphs[i].n entries. */
int smdp_sssp_multi_graph(smdp_ctx* ctx, const smdp_csr* graphs, int32_t ng, const int32_t* srcs,
                          int32_t* const* dist, smdp_stats* st);

/* ---- resident graphs: upload once, run many times -------------------------*/
int  smdp_graph_create(smdp_ctx* ctx, const smdp_csr* g, smdp_graph** out);
void smdp_graph_destroy(smdp_graph* gr);
/* SpMV values (m) and x (n); PageRank out-degrees (n); coloring priorities (n). */
int  smdp_graph_set_values(smdp_graph* gr, const float* val);
int  smdp_graph_set_x(smdp_graph* gr, const float* x);
int  smdp_graph_set_outdeg(smdp_graph* gr, const int32_t* outdeg);
int  smdp_graph_set_prio(smdp_graph* gr, const uint32_t* prio);

int  smdp_sssp_run(smdp_graph* gr, int32_t src, int solution, smdp_stats* st);
int  smdp_bfs_run(smdp_graph* gr, int32_t src, int solution, smdp_stats* st);
int  smdp_spmv_run(smdp_graph* gr, int solution, smdp_stats* st);
int  smdp_pagerank_run(smdp_graph* gr, const smdp_pr_opts* opt, int solution, smdp_stats* st);
int  smdp_color_run(smdp_graph* gr, int use_prio, int solution, smdp_stats* st);
int  smdp_sssp_multi_source_run(smdp_graph* gr, const int32_t* srcs, int32_t k, smdp_stats* st);
/* smdp_tree on a resident graph; results: SMDP_RES_HEIGHT and SMDP_RES_DESC. */
int  smdp_tree_run(smdp_graph* gr, int32_t root, smdp_stats* st);

/* Copy the last result of the named kind to the host. */
#define SMDP_RES_DIST   0   /* int32[n]   SSSP */
#define SMDP_RES_LEVEL  1   /* int32[n]   BFS */
#define SMDP_RES_Y      2   /* float[n]   SpMV */
#define SMDP_RES_RANK   3   /* float[n]   PageRank */
#define SMDP_RES_COLOR  4   /* int32[n]   coloring */
#define SMDP_RES_MDIST  5   /* int32[k*n] multi-source SSSP, source-major */
#define SMDP_RES_HEIGHT 6   /* int32[n]   tree: subtree heights */
#define SMDP_RES_DESC   7   /* int32[n]   tree: descendant counts (same run as SMDP_RES_HEIGHT) */
int  smdp_graph_read(smdp_graph* gr, int what, void* host_out, int64_t count);

/* Batched small-graph SSSP kept resident: upload once, replay the captured
 * hipGraph many times. */
typedef struct smdp_batch smdp_batch;
int  smdp_batch_create(smdp_ctx* ctx, const smdp_csr* graphs, int32_t ng, const int32_t* srcs,
                       smdp_batch** out);
int  smdp_batch_run(smdp_batch* b, smdp_stats* st);
int  smdp_batch_read(smdp_batch* b, int32_t i, int32_t* dist);
void smdp_batch_destroy(smdp_batch* b);
/* Times the batch under candidate plans (workgroups per CU 1 or 2, group_share
 * 50, 100 or 200; median of five replays after a warm-up), keeps the fastest
 * for this batch and installs it in the context's SMDP_APP_MSSSP parameters,
 * so later batches start from it.  No reference counterpart: the reference
 * solves one graph per call. */
int  smdp_batch_autotune(smdp_batch* b, smdp_params* chosen);

/* ---- autotuner --------------------------------------------------------------*/
/* Times candidate parameter sets of SMDP_SOL_TUNED for `app` on the resident
 * graph and installs the fastest in the context.  cache_path (may be NULL) is a
 * JSON file keyed by GPU arch and the graph's degree histogram; a hit skips the
 * search.  For SSSP/BFS `src` is the probe source. */
int smdp_autotune(smdp_graph* gr, int app, int32_t src, const char* cache_path,
                  smdp_params* chosen);

/* ---- an entry point that allocates nothing --------------------------------*/
/* smdp_sssp (same solutions, same distances) on a CSR the caller already holds
 * in device memory: d_row[n+1], d_col[m], d_w[m] and the output d_dist[n] are
 * DEVICE pointers owned by the caller.  No counterpart in the reference; it is
 * for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The call allocates
 * nothing: every working buffer (frontier queues, stamps, grid queue, control
 * blocks, counters) comes from one fixed 12 MB device array that the library
 * holds on each GPU it runs on until the process ends (set up when the first
 * context on that GPU is created), so n <= SMDP_NOALLOC_MAX_N and
 * m <= SMDP_NOALLOC_MAX_M, else SMDP_ERR_ARG.  The CSR is checked on the device
 * first (SMDP_ERR_ARG when row is not monotone from 0 to m or a column is out of
 * range).  All calls on one GPU share that array, so calls that overlap on one
 * GPU (from other threads or contexts) race.  Profiling is ignored, and
 * SMDP_SOL_TUNED runs unordered rounds on the graph as given (no renumbered
 * copy, no buckets). */
#define SMDP_NOALLOC_MAX_N (1 << 19)
#define SMDP_NOALLOC_MAX_M (1 << 23)
int smdp_sssp_device_noalloc(smdp_ctx* ctx, int32_t n, int32_t m, const int32_t* d_row, const int32_t* d_col,
                             const int32_t* d_w, int32_t src, int solution, int32_t* d_dist, smdp_stats* st);

/* ---- host graph utilities (no device needed) ------------------------------*/
/* A host graph owned by the library. */
typedef struct {
  int32_t  n;
  int32_t  m;
  int32_t* row;      /* n+1 */
  int32_t* col;      /* m */
  int32_t* w;        /* m or NULL */
  int32_t  source;   /* 0-based source from the file ('s' line) or generator */
} smdp_hgraph;

#define SMDP_FMT_DIMACS9  0   /* common/global.cpp:29-64 */
#define SMDP_FMT_DIMACS10 1   /* common/global.cpp:66-105 */
#define SMDP_FMT_SLNDC    2   /* common/global.cpp:107-190 */

/* readInputDIMACS9/10/SLNDC + convertCSR: CSR with each row in reverse file
 * order, as convertCSR builds it (common/global.cpp:192-235). */
int  smdp_read_graph(const char* path, int format, smdp_hgraph** out);
/* invertGraph (common/global.cpp:237-277): in-CSR, each row in descending
 * source order. */
int  smdp_invert(const smdp_hgraph* g, smdp_hgraph** out);
/* Write DIMACS9 text ('p sp', 's', 'a' lines, 1-based). */
int  smdp_write_dimacs9(const char* path, const smdp_hgraph* g);
void smdp_hgraph_free(smdp_hgraph* g);

/* Seeded generators.  Every random draw is SplitMix64 of a counter
 * (seed*2^40 + stream*2^32 + i), so outputs do not depend on thread count. */
/* n vertices, deg uniform out-arcs each, weights in [wlo,whi], source 0. */
int smdp_gen_uniform(int32_t n, int32_t deg, int32_t wlo, int32_t whi, uint64_t seed,
                     smdp_hgraph** out);
/* RMAT 2^scale vertices, ef*2^scale arcs, quadrant probabilities a,b,c
 * (d = 1-a-b-c); weights in [wlo,whi] (whi == 0: no weights); symmetrize adds
 * the reverse of every arc; permute relabels vertices by a seeded permutation.
 * source = a seeded random vertex with out-degree > 0. */
int smdp_gen_rmat(int32_t scale, int32_t ef, double a, double b, double c, int32_t wlo,
                  int32_t whi, int symmetrize, int permute, uint64_t seed, smdp_hgraph** out);
/* Rooted tree in BFS order: node i gets k children, k ~ Zipf(s) on [1,cap],
 * until n nodes exist.  Arcs parent -> child, root 0. */
int smdp_gen_zipf_tree(int32_t n, double s, int32_t cap, uint64_t seed, smdp_hgraph** out);
/* Graph i of a small-graph batch: V = round(exp(U[ln vlo, ln vhi])),
 * M = round(V * U[dlo, dhi]) uniform arcs, weights in [wlo,whi], source 0. */
int smdp_gen_small_graph(int32_t i, int32_t vlo, int32_t vhi, double dlo, double dhi,
                         int32_t wlo, int32_t whi, uint64_t seed, smdp_hgraph** out);
/* k distinct seeded random sources with out-degree > 0. */
int smdp_gen_sources(const smdp_hgraph* g, int32_t k, uint64_t seed, int32_t* srcs);
/* Uniform FP32 in [lo,hi) for i in [0,count): values for SpMV. */
int smdp_gen_uniform_f32(int64_t count, float lo, float hi, uint64_t seed, uint32_t stream,
                         float* out);
/* Uniform u32 priorities. */
int smdp_gen_prio(int32_t n, uint64_t seed, uint32_t* out);
/* The reference SpMV inputs: data[i] = i%13+0.5, x[i] = i%17+0.5
 * (apps/SpMV/src/SpMV_wrapper.cu:105-110). */
int smdp_gen_spmv_reference_inputs(int32_t n, int32_t m, float* data, float* x);
/* Out-degrees of a CSR (pagerank.cpp:108-113). */
int smdp_outdegree(const smdp_hgraph* g, int32_t* outdeg);

#ifdef __cplusplus
}
#endif
#endif /* SMDP_H */
