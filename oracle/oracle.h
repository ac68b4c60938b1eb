/*
 * oracle.h -- CPU restatement of the reference's CPU checkers for the GAS
 * frontier path (BFS, SSSP, CC, PageRank) plus FP64 PageRank / adjoint
 * solvers.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU baseline).  The product library (libgasgraph.so) never
 * links or calls it.
 *
 * Parity pinning (each function is checked in tests/test_oracle.py against an
 * independent source -- hand-derived KATs, scipy.sparse.csgraph, dense FP64
 * linear solves, finite differences -- never against a retyped formula):
 *   oracle_bfs                 pinned: KATs + scipy BFS + |d[u]-d[v]|<=1
 *   oracle_bfs_parents         pinned: hand-derived KAT parents + accepted by
 *                              oracle_bfs_validate + brute-force min rule
 *   oracle_bfs_validate        pinned: accepts scipy's BFS tree (any valid
 *                              parents) and rejects one planted error per rule
 *   oracle_sssp_selection      pinned: KATs + scipy dijkstra
 *   oracle_sssp_dijkstra       pinned: equal to oracle_sssp_selection + scipy
 *   oracle_cc                  pinned: scipy connected_components (min id)
 *   oracle_pr_reference        pinned: dense solve of its fixed point (tol)
 *   oracle_pagerank_f64        pinned: dense FP64 linear solve + sum(r)=1
 *   oracle_pagerank_adjoint_f64 pinned: finite differences + 3-vertex hand case
 *   oracle_bfs_omp, oracle_sssp_omp  pinned: equal to oracle_bfs / oracle_sssp
 *                              (the parallel CPU baselines)
 */
#ifndef GG_ORACLE_H
#define GG_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Algorithms/BFS/bfs.cu:57-111 (CPUBFS). depth[v] = -1 if unreached. */
void oracle_bfs(int32_t n, const int32_t *ro, const int32_t *col, int32_t src,
                int32_t *depth);

/* BFS tree parents: pred[src] = src, pred[v] = smallest in-neighbour one
 * level up, -1 unreached.  Same rule as gg_bfs_tree. */
void oracle_bfs_parents(int32_t n, const int32_t *ro, const int32_t *col,
                        const int32_t *depth, int32_t src, int32_t *pred);

/* Graph500 BFS validation (Algorithms/BFS/bfs.cu:984-1012): 0 = valid, else
 * the number of the first rule broken (see oracle.c). */
int32_t oracle_bfs_validate(int32_t n, const int32_t *ro, const int32_t *col,
                            int32_t src, const int32_t *depth,
                            const int32_t *pred);

/* Algorithms/SSSP/sssp.cu:126-183 (CPUSSSP) as written: O(n^2) selection,
 * sentinel 100000000, sums formed in 64-bit ("long"). */
void oracle_sssp_selection(int32_t n, const int32_t *ro, const int32_t *col,
                           const int32_t *w, int32_t src, int32_t *dist);

/* Same output contract as oracle_sssp_selection, binary-heap Dijkstra
 * (O(m log n)) for graphs where the O(n^2) form is infeasible. */
void oracle_sssp_dijkstra(int32_t n, const int32_t *ro, const int32_t *col,
                          const int32_t *w, int32_t src, int32_t *dist);

/* Algorithms/CC/cc.cu:127-169 (CPUCC): in-place min-label sweeps. Returns
 * the number of sweeps. */
int32_t oracle_cc(int32_t n, const int32_t *ro, const int32_t *col,
                  int32_t *label);

/* Algorithms/PageRank/PageRank.cu:108-161 (CPUPR): in-place Gauss-Seidel,
 * r = 0.15 + 0.85 * sum_{u in in(v)} r_u / outdeg_u, stop when no vertex
 * moves by >= 0.01.  outdeg = CSR out-degree (adjacent_difference of
 * row_offsets).  co/ri = CSC.  Returns the sweep count. */
int32_t oracle_pr_reference(int32_t n, const int32_t *ro, const int32_t *co,
                            const int32_t *ri, float *rank);

/* North-star PageRank in FP64 (Jacobi power iteration):
 *   r = d * (P^T r + (sum_{u dangling} r_u) / n * 1) + (1-d) * t
 * P_uv = p[e] for CSR edge e=(u,v) (p == NULL: 1/outdeg(u)); t == NULL means
 * uniform 1/n.  Stops when ||r_new - r||_1 <= tol or after max_iter sweeps.
 * Returns the sweep count, writes the last residual to *resid. */
int32_t oracle_pagerank_f64(int32_t n, const int32_t *ro, const int32_t *col,
                            const double *p, const double *t, double d,
                            double tol, int32_t max_iter, double *r,
                            double *resid);

/* FP64 adjoint of oracle_pagerank_f64 for L = sum_v c_v r_v.
 *   lambda = c + d * (P lambda + delta * (1^T lambda) / n)   (delta_u = dangling)
 *   dL/dt  = (1-d) lambda
 *   dL/dw_e (e=(u,v), P_uv = w_uv / S_u) = d r_u / S_u * (lambda_v - sum_v' P_uv' lambda_v')
 * w == NULL means P = 1/outdeg (then grad_w is dL/dP_uv = d r_u lambda_v).
 * Returns sweep count. */
int32_t oracle_pagerank_adjoint_f64(int32_t n, const int32_t *ro,
                                    const int32_t *col, const double *w,
                                    double d, const double *r, const double *c,
                                    double tol, int32_t max_iter,
                                    double *lambda, double *grad_w,
                                    double *grad_t);

/* Central finite differences of L(w) = sum_v c_v r_v(w) (r as in
 * oracle_pagerank_f64 with P_uv = w_uv / S_u, any teleport) for `ne` CSR
 * edges: fd[i] = (L(w + h_i e_i) - L(w - h_i e_i)) / (2 h_i).  `r` is the
 * FP64 fixed point at w.  Each perturbed fixed point is r + D, where D solves
 * the difference of the two fixed-point equations (only row u of P changes):
 *   D = d (P'^T D + (delta . D) / n 1) + d r_u (P'_u. - P_u.)
 * iterated from 0 until every ||D_new - D||_1 <= tol_rel ||D_new||_1, so
 * L(w + h) - L(w - h) = c . (D+ - D-) carries no cancellation against L.
 * Up to 32 edges (64 perturbations) share a sweep (OpenMP).  Returns the
 * largest sweep count. */
int32_t oracle_pagerank_fd_f64(int32_t n, const int32_t *ro, const int32_t *col,
                               const double *w, double d, const double *r,
                               const double *c, int32_t ne, const int64_t *edges,
                               const double *h, double tol_rel, int32_t max_iter,
                               double *fd);

/* Parallel CPU baselines (OpenMP, oracle_threads() threads): outputs equal
 * to oracle_bfs / oracle_sssp_selection.  in_ro / in_col: in-edges (the
 * CSR itself for a symmetric graph); delta: bucket width (e.g. 128). */
int32_t oracle_threads(void);
void oracle_bfs_omp(int32_t n, const int32_t *ro, const int32_t *col,
                    const int32_t *in_ro, const int32_t *in_col, int32_t src,
                    int32_t *depth);
void oracle_sssp_omp(int32_t n, const int32_t *ro, const int32_t *col,
                     const int32_t *w, int32_t src, int32_t delta,
                     int32_t *dist);

/* Min-label CC on all host cores: in-place sweeps as oracle_cc, each thread
 * updating its own static chunk of vertices and reading the others' labels
 * with relaxed atomic loads (labels only decrease, so a stale read only
 * defers an update to the next sweep); same unique fixed point (the
 * smallest id of each component).  Returns the sweep count. */
int32_t oracle_cc_omp(int32_t n, const int32_t *ro, const int32_t *col,
                      int32_t *label);

/* oracle_pagerank_f64 on all host cores, as a pull over in-edges: co / ri =
 * CSC, p_csc = transition weight per CSC slot (NULL: 1 / outdeg(u)), outdeg
 * from ro.  Same iteration (Jacobi), same stopping rule; sums in another
 * order, so equal to the serial result within rounding. */
int32_t oracle_pagerank_f64_omp(int32_t n, const int32_t *ro, const int32_t *co,
                                const int32_t *ri, const double *p_csc,
                                const double *t, double d, double tol,
                                int32_t max_iter, double *r, double *resid);

#ifdef __cplusplus
}
#endif
#endif
