/*
 * gasgraph.h -- C ABI of libgasgraph.so, the MI355X (gfx950) Gather-Apply-
 * Scatter frontier engine.
 *
 * The reference (MapGraph, a header-only CUDA C++ engine) has no FFI: callers
 * instantiate EnactorVertexCentric<CsrProblem, Program> in their own
 * translation unit.  Each entry point below replaces one step of that
 * sequence for the four shipped vertex programs; the citation on each names
 * the reference interface it stands in for.  The same engine is available as
 * header-only HIP templates (gpugraph_amd/csrc/gasengine/) for user programs.
 *
 * Conventions
 *  - Every call returns int status: GG_OK (0) or a GG_ERR_* code; the message
 *    of the last failure on the calling thread is gg_last_error().  The
 *    library never exits or aborts (the reference calls exit()/abort(),
 *    GASengine/enactor_vertex_centric.cuh:107-116,622-623,768-772).
 *  - Vertex ids, row offsets and column indices are int32 (m < 2^31).
 *  - Output pointers are host memory unless GG_OUT_DEVICE is set, in which
 *    case they are device pointers on the graph's device and no copy is made.
 *  - All work of one graph handle is issued on its stream (gg_graph_set_stream);
 *    calls on distinct handles are independent.  One handle must not be used
 *    from two threads at once.
 */
#ifndef GASGRAPH_H
#define GASGRAPH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define GG_OK 0
#define GG_ERR_INVALID 1        /* bad argument (null pointer, id out of range, ...) */
#define GG_ERR_OOM 2            /* device or host allocation failed */
#define GG_ERR_OVERFLOW 3       /* a queue or capacity limit was exceeded */
#define GG_ERR_HIP 4            /* a HIP runtime call failed */
#define GG_ERR_NOT_CONVERGED 5  /* iterative solver hit max_iter above tol */
#define GG_ERR_NO_DEVICE 6      /* no HIP device visible */
#define GG_ERR_IO 7             /* file open / parse failure */
#define GG_ERR_UNSUPPORTED 8    /* option not supported for this graph */
#define GG_ERR_INTERNAL 9       /* an internal consistency check failed (a library bug) */

/* ---- edge value dtypes ------------------------------------------------- */
#define GG_VALUES_NONE 0
#define GG_VALUES_I32 1
#define GG_VALUES_F32 2

/* ---- graph flags ------------------------------------------------------- */
#define GG_GRAPH_DIRECTED 1u    /* build the CSC (in-edges) on the device;
                                   otherwise the CSR must be symmetric and is
                                   used as its own CSC */
#define GG_GRAPH_DEVICE_PTRS 2u /* inputs are device pointers on the current device */
#define GG_GRAPH_REORDER 4u     /* store the graph renumbered by out-degree, highest first (locality
                                   of the random vertex gathers); every call still takes and
                                   returns the caller's vertex ids and edge order */

/* ---- run flags --------------------------------------------------------- */
#define GG_OUT_DEVICE 1u        /* output pointers are device memory */
#define GG_INSTRUMENT 2u        /* record per-kernel HIP-event timings (gg_profile_*) */
#define GG_BFS_PUSH_ONLY 4u     /* BFS through the generic GAS program path (no bottom-up) */
#define GG_CC_HOOK 8u           /* CC by min-id hooking + pointer jumping instead of label propagation */
#define GG_SSSP_GAS 16u         /* SSSP through the generic GAS program path (reference semantics) */
#define GG_CC_GAS 32u           /* CC label propagation through the generic GAS program path */
#define GG_BFS_REVERSE 64u      /* BFS over in-edges (hops to src), GAS path, EXPAND_IN_EDGES */
#define GG_CC_WEAK 128u         /* directed graphs: weakly connected components, GAS path,
                                   EXPAND_ALL_EDGES */
#define GG_PR_ALL_EDGES 256u    /* gg_pagerank_reference on a directed graph read as undirected:
                                   GATHER_ALL_EDGES, degree = out + in */

typedef struct gg_graph gg_graph;

/* Engine tuning knobs.  The reference hard-codes the degree tiers (32/512,
 * enactor_vertex_centric.cuh:3949-3950) and takes the LBS/dynamic switch from
 * a per-algorithm "threshold" parameter (Algorithms/SSSP/register.h:39). */
typedef struct gg_options {
  int32_t threshold;     /* |frontier| above which the scan-based LBS expand is used; -1: the
                            algorithm's default, and then a smaller frontier holding a row
                            longer than 4 * block_degree is expanded by LBS as well */
  int32_t wave_degree;   /* degree >= this: one wave64 per vertex (tiered expand; also the
                            longest in-edge list a single lane scans in bottom-up BFS) */
  int32_t block_degree;  /* degree >= this: whole workgroup per vertex (tiered expand) */
  int32_t iter_num;      /* maximum GAS iterations (reference "iter_num") */
  int32_t bfs_alpha;     /* top-down -> bottom-up when m_f > m_u / alpha */
  int32_t bfs_beta;      /* bottom-up -> top-down when n_f < n / beta */
  int32_t lbs_tile;      /* merge-path positions per LBS workgroup tile: 512, 1024 or 2048 */
  int32_t sssp_delta;    /* SSSP bucket width (0 = auto: 128); INT32_MAX = one bucket */
  int32_t pull_alpha;    /* SSSP: pull step when frontier edges > m / pull_alpha (0 = never) */
  int32_t expand_grid;   /* persistent workgroups of the LBS expand kernel (64 .. 65536) */
  int32_t xcd_slices;    /* PageRank sweeps split by gathered-vertex slice, one slice per
                            XCD L2: -1 = auto (currently off), 0 = off, 1 = on */
  int32_t gas_expand_grid; /* workgroups of the generic GAS path's LBS expand and gather
                              kernels (64 .. 65536): their per-edge work is the program's
                              functors, longer chains than the specialised kernels' */
} gg_options;

/* Per-run statistics (the reference prints these with INSTRUMENT/verbose). */
typedef struct gg_stats {
  int32_t iterations;       /* GAS iterations (BFS levels) */
  int32_t pull_iterations;  /* of which bottom-up / pull steps */
  int64_t reached_vertices; /* vertices with a final value != sentinel */
  int64_t reached_edges;    /* sum of CSR out-degree over reached vertices (GTEPS numerator) */
  int64_t edges_inspected;  /* edges actually examined, summed over iterations */
  double alg_bytes;         /* algorithmic HBM bytes of the run (byte model of the kernel set) */
  double gpu_ms;            /* HIP-event time of the device section */
  double residual;          /* PageRank: last L1 residual (FP64); else 0 */
} gg_stats;

/* ---- library ------------------------------------------------------------ */
const char *gg_last_error(void);
int gg_version(void);                     /* major*10000 + minor*100 + patch */
int gg_device_count(int32_t *count);
int gg_default_options(gg_options *out);

/* ---- graph -------------------------------------------------------------- */
/* Replaces CsrProblem::FromHostProblem (GASengine/csr_problem.cuh:512-839):
 * copies the CSR to the current device and, for GG_GRAPH_DIRECTED, builds the
 * CSC and the CSC->CSR edge-id permutation (reference: thrust stable sort in
 * util.cuh:55-126, called csr_problem.cuh:754-765). values may be NULL. */
int gg_graph_create(int32_t n, int64_t m, const int32_t *row_offsets,
                    const int32_t *col_indices, const void *values,
                    int32_t value_dtype, uint32_t flags, gg_graph **out);
/* Replaces GraphSlice::~GraphSlice (csr_problem.cuh:234-353). */
int gg_graph_destroy(gg_graph *g);
/* stream: a hipStream_t on the graph's device, or NULL for the library's own. */
int gg_graph_set_stream(gg_graph *g, void *stream);
int gg_graph_set_options(gg_graph *g, const gg_options *opt);
int gg_graph_get_options(const gg_graph *g, gg_options *opt);
/* Replace the edge values (same dtype, m entries) without rebuilding the
 * topology; host pointer unless GG_OUT_DEVICE is set in flags. */
int gg_graph_update_values(gg_graph *g, const void *values, uint32_t flags);
int gg_graph_info(const gg_graph *g, int32_t *n, int64_t *m, int32_t *directed,
                  int32_t *value_dtype);
/* Device pointers of the resident CSR (for zero-copy callers). */
int gg_graph_device_csr(const gg_graph *g, const int32_t **row_offsets,
                        const int32_t **col_indices, const void **values);
/* Device pointers of the in-edge view: for directed graphs the CSC built on
 * the device at creation and its CSC->CSR edge-id permutation (reference:
 * d_column_offsets / d_row_indices / d_edge_id of GraphSlice, built by
 * util.cuh:55-126); for undirected graphs the CSR itself and NULL ids.
 * Within a column, source rows ascend. */
int gg_graph_device_csc(const gg_graph *g, const int32_t **column_offsets,
                        const int32_t **row_indices, const int32_t **csc_edge_ids);
/* GG_GRAPH_REORDER graphs: device arrays new_of_old[v] (internal id of
 * caller vertex v) and old_of_new[i]; the device CSR / CSC views above are in
 * internal numbering.  Both NULL for graphs stored in caller order. */
int gg_graph_vertex_order(const gg_graph *g, const int32_t **new_of_old,
                          const int32_t **old_of_new);
/* Wait for all work issued on the graph's stream. */
int gg_graph_synchronize(gg_graph *g);

/* ---- vertex programs ---------------------------------------------------- */
/* BFS depths (src = 0, unreached = -1).  Replaces EnactIterativeSearch with
 * Program = bfs (Algorithms/BFS/bfs.h:28-298; driver bfs.cu:873-881) plus
 * ExtractResults (csr_problem.cuh:433-458).  Default: direction-optimising
 * (top-down LBS push / bottom-up bitmap pull); GG_BFS_PUSH_ONLY runs the
 * generic GAS program path. */
int gg_bfs(gg_graph *g, int32_t src, int32_t *depth, uint32_t flags,
           gg_stats *stats);

/* BFS depths and a BFS tree.  Replaces the reference's predecessor marking
 * (CsrProblem<..., MARK_PREDECESSORS> csr_problem.cuh:113-117, extraction
 * Algorithms/BFS/predextract.cuh:246+): after the BFS of gg_bfs (same flags),
 * one pass picks for every reached vertex v != src its smallest in-neighbour
 * one level up, so the tree is unique; pred[src] = src, pred = -1 where
 * unreached.  depth may be NULL.  The tree passes the reference's validation
 * rules (Algorithms/BFS/bfs.cu:984-1012). */
int gg_bfs_tree(gg_graph *g, int32_t src, int32_t *depth, int32_t *pred,
                uint32_t flags, gg_stats *stats);

/* SSSP distances with int32 edge weights (unreached = 100000000).  Replaces
 * EnactIterativeSearch with Program = sssp (Algorithms/SSSP/sssp.h:34-393;
 * driver sssp.cu:389-404). */
int gg_sssp(gg_graph *g, int32_t src, int32_t *dist, uint32_t flags,
            gg_stats *stats);

/* Connected-component labels = min vertex id of the component.  Replaces
 * EnactIterativeSearch with Program = cc (Algorithms/CC/cc.h:32-349; driver
 * cc.cu:291-292).  Default: the program's label propagation as a
 * device-driven loop (hipGraph replays, one host check per 16 iterations);
 * GG_CC_GAS: the same program through the generic GAS path; GG_CC_HOOK:
 * hooking + pointer jumping.  Undirected graphs. */
int gg_cc(gg_graph *g, int32_t *label, uint32_t flags, gg_stats *stats);

/* The reference's PageRank program (Algorithms/PageRank/PageRank.h:33-409):
 * r = 0.15 + 0.85 * sum r_u/outdeg_u over in-edges, a vertex stays active
 * while |delta| >= 0.01; FP32 through the generic GAS gather/apply path. */
int gg_pagerank_reference(gg_graph *g, float *rank, uint32_t flags,
                          gg_stats *stats);

/* PageRank r = d (P^T r + dangling/n) + (1-d) t, P_uv = w_uv / sum_v' w_uv'
 * (F32 values) or 1/outdeg(u) (no values, or int32 values, which are SSSP
 * weights and do not enter PageRank).  FP32 gather with Kahan
 * compensation, FP64 dangling sum and L1 residual; stops when the L1
 * residual <= tol_l1 (GG_ERR_NOT_CONVERGED after max_iter sweeps).
 * teleport: n floats or NULL for uniform.  No reference counterpart beyond
 * the pagerank program's gather/apply (PageRank.h:197-281). */
int gg_pagerank(gg_graph *g, float damping, const float *teleport,
                double tol_l1, int32_t max_iter, float *rank, uint32_t flags,
                gg_stats *stats);

/* Adjoint of gg_pagerank for L = sum_v c_v r_v, given the stored forward
 * ranks: grad_w (m, CSR order) = dL/dw for the raw F32 values (or dL/dP_uv
 * when the graph has no F32 values); grad_t (n) = dL/dt.  Either output may
 * be NULL.  The adjoint iteration stops when its L1 residual is at most
 * tol_l1 * ||c||_1.  With GG_OUT_DEVICE every array argument (rank, c and
 * the outputs) is a device pointer.  No reference counterpart. */
int gg_pagerank_adjoint(gg_graph *g, float damping, const float *rank,
                        const float *c, double tol_l1, int32_t max_iter,
                        float *grad_w, float *grad_t, uint32_t flags,
                        gg_stats *stats);

/* ---- batched instances -------------------------------------------------- */
/* num_sssp SSSP sources and num_ppr personalized-PageRank seeds (one-hot
 * teleports) solved together: per-instance frontiers in shared kernels,
 * iteration loop captured as a HIP graph; num_bfs BFS sources, 64 per pass,
 * one bit per instance in a 64-bit frontier word per vertex.  Outputs are
 * instance-major: sssp_dist[k*n + v], ppr_rank[k*n + v], bfs_depth[k*n + v]
 * (the BFS program's depths, -1 unreached).  Replaces the drivers'
 * sequential per-source loops (sssp.cu:389-404, bfs.cu:865-881).  Fields
 * after max_iter were added later; zero-initialise the struct. */
typedef struct gg_batch_spec {
  int32_t num_sssp;
  const int32_t *sssp_sources;
  int32_t *sssp_dist;
  int32_t num_ppr;
  const int32_t *ppr_seeds;
  float *ppr_rank;
  float damping;
  double tol_l1;
  int32_t max_iter;
  int32_t num_bfs;
  const int32_t *bfs_sources;
  int32_t *bfs_depth;
} gg_batch_spec;

int gg_batch_run(gg_graph *g, const gg_batch_spec *spec, uint32_t flags,
                 gg_stats *stats);

/* ---- autotuner ---------------------------------------------------------- */
/* Replaces the reference's fixed schedule constants (degree tiers 32/512,
 * enactor_vertex_centric.cuh:3949-3950) and hand-set "threshold"
 * (Algorithms/SSSP/register.h:39): coordinate descent over the gg_options
 * knobs on this graph, timing the selected programs from `src` (fastest of
 * `reps` runs, each candidate bracketed by runs of the current options) and
 * keeping a change when the programs it affects get >= 3% faster; a final
 * bracketed run against the starting options keeps the result only if it
 * still wins.  Every trial's outputs must equal the first run's (device
 * checksum) or the call fails with GG_ERR_INTERNAL.  On success the best
 * options are installed on the graph and copied to *best; trials[] (up to
 * max_trials, count in *num_trials) logs each timing: first the untuned
 * baseline (option -1), last the final comparison (option -2) when any knob
 * changed.  On failure the graph keeps the options it had. */
#define GG_TUNE_BFS 1u   /* gg_bfs, direction-optimising */
#define GG_TUNE_SSSP 2u  /* gg_sssp, near-far (int32 values only) */
#define GG_TUNE_CC 4u    /* gg_cc, label propagation (undirected only) */
#define GG_TUNE_GAS 8u   /* reference-form programs: gg_bfs GG_BFS_PUSH_ONLY, gg_sssp GG_SSSP_GAS */
#define GG_TUNE_PAGERANK 16u /* gg_pagerank, damping 0.85, uniform teleport, tol 1e-6 (FP32
                                results: a trial must stay within L1 1e-5 of the baseline) */

typedef struct gg_tune_trial {
  uint32_t programs;  /* GG_TUNE_* programs timed in this trial */
  int32_t option;     /* index of the int32 field of gg_options varied (0 = threshold ...
                         10 = xcd_slices); -1 = baseline */
  int32_t value;      /* the value tried */
  float ms;           /* summed median device time of those programs */
} gg_tune_trial;

int gg_autotune(gg_graph *g, uint32_t programs, int32_t src, int32_t reps,
                gg_options *best, gg_tune_trial *trials, int32_t max_trials,
                int32_t *num_trials);

/* ---- instrumentation ---------------------------------------------------- */
/* Per-kernel-class event timings collected by runs with GG_INSTRUMENT
 * (the reference's INSTRUMENT template flag, enactor_vertex_centric.cuh:96). */
typedef struct gg_kernel_stat {
  char name[48];
  int64_t launches;
  double total_ms;
  double alg_bytes;  /* algorithmic bytes attributed to this kernel class */
} gg_kernel_stat;

int gg_profile_reset(gg_graph *g);
int gg_profile_get(gg_graph *g, gg_kernel_stat *out, int32_t max_entries,
                   int32_t *count);

/* The same timings launch by launch (oldest first, at most 65536 kept until
 * gg_profile_reset), for per-iteration bandwidth: alg_bytes is the launch's
 * own byte-model share where the run attributes it per launch (BFS and SSSP
 * frontier kernels), else 0. */
typedef struct gg_launch_stat {
  char name[48];
  double ms;
  double alg_bytes;
} gg_launch_stat;

int gg_profile_launches(gg_graph *g, gg_launch_stat *out, int32_t max_entries,
                        int32_t *count);

/* ---- allocation-free entry point ---------------------------------------- */
/* BFS over a CSR already on the device, allocating nothing: d_row_offsets
 * (n+1) and d_col_indices (m) in device memory, depths (the BFS program's:
 * 0 at src, -1 unreached, same as gg_bfs) written to d_depth (n, device).
 * Scratch (queues, visited bitmap, scan and partition arrays, counters) is a
 * fixed device array of the library, about 25 MB per GPU, held until the
 * process ends; n <= 2^21 and m <= 2^26, larger graphs get GG_ERR_INVALID.
 * Work runs on `stream` (a hipStream_t; NULL = the null stream) with one
 * synchronisation per level.  Calls that overlap on one GPU share that
 * scratch and race.  No counterpart in the reference: for callers that
 * cannot allocate (serving loops, allocation-free harnesses). */
int gg_bfs_noalloc(int32_t n, int64_t m, const int32_t *d_row_offsets,
                   const int32_t *d_col_indices, int32_t src, int32_t *d_depth,
                   void *stream, gg_stats *stats);

/* ---- graph construction (host) ----------------------------------------- */
/* Host CSR produced by the builders below; free with gg_csr_free. */
typedef struct gg_csr {
  int32_t n;
  int64_t m;
  int32_t *row_offsets;  /* n+1 */
  int32_t *col_indices;  /* m */
  void *values;          /* m entries of value_dtype, or NULL */
  int32_t value_dtype;
} gg_csr;

void gg_csr_free(gg_csr *csr);

/* Matrix Market reader (b40c/graph/builder/market.cuh:42-209 +
 * csr_graph.cuh FromCoo:121-150): 1-based ids, optional mirror when
 * undirected, rows sorted by (row, col) with a stable sort, duplicates and
 * self-loops kept.  "pattern" files get value 1. */
int gg_read_mtx(const char *path, int32_t undirected, int32_t value_dtype,
                gg_csr *out);

/* Seeded generators.  Same (parameters, seed) -> same CSR on every host. */
/* G(n, M): M distinct undirected pairs u != v, symmetric CSR. */
int gg_gen_erdos_renyi(int32_t n, int64_t num_pairs, uint64_t seed, gg_csr *out);
/* R-MAT(a,b,c) with 2^scale vertices and edge_factor*2^scale draws; vertex
 * ids scrambled by a seeded bijection; self-loops and duplicates removed;
 * symmetric when undirected. */
int gg_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b,
                double c, int32_t undirected, uint64_t seed, gg_csr *out);
/* side x side 4-neighbour lattice, each bond kept with probability keep_p,
 * plus floor(long_range_frac * kept) random long-range pairs; symmetric. */
int gg_gen_lattice(int32_t side, double keep_p, double long_range_frac,
                   uint64_t seed, gg_csr *out);
/* Attach values to a CSR: int32 uniform in [lo, hi] (symmetric in (u,v) when
 * symmetric != 0), or F32 (k+1)/2^24 with k 24 random bits (lo/hi ignored). */
int gg_csr_weights(gg_csr *csr, int32_t value_dtype, int32_t lo, int32_t hi,
                   int32_t symmetric, uint64_t seed);

#ifdef __cplusplus
}
#endif
#endif /* GASGRAPH_H */
