/*
 * oracle.c -- CPU restatement of the reference's CPU checkers (see oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  Never linked into or called by the product
 * library.  Plain C99, single-threaded, so its timings are a 1-core baseline.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define SSSP_INF 100000000 /* Algorithms/SSSP/sssp.h:43 INIT_VALUE */

/* Algorithms/BFS/bfs.cu:57-111 -- FIFO queue BFS, depth = parent depth + 1. */
void oracle_bfs(int32_t n, const int32_t *ro, const int32_t *col, int32_t src,
                int32_t *depth) {
  for (int32_t i = 0; i < n; i++) depth[i] = -1;
  if (src < 0 || src >= n) return;
  int32_t *queue = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  int32_t head = 0, tail = 0;
  depth[src] = 0;
  queue[tail++] = src;
  while (head < tail) {
    int32_t u = queue[head++];
    int32_t nd = depth[u] + 1;
    for (int32_t e = ro[u]; e < ro[u + 1]; e++) {
      int32_t v = col[e];
      if (depth[v] == -1) {
        depth[v] = nd;
        queue[tail++] = v;
      }
    }
  }
  free(queue);
}

/* BFS tree parents in the form the library returns: pred[src] = src,
 * pred[v] = the smallest u with an edge u->v and depth[u] = depth[v] - 1
 * for every other reached v, -1 for unreached v.  (The reference marks the
 * discovering vertex, MARK_PREDECESSORS of csr_problem.cuh:113-117; any
 * valid parent passes its validation, this rule makes the choice unique.) */
void oracle_bfs_parents(int32_t n, const int32_t *ro, const int32_t *col,
                        const int32_t *depth, int32_t src, int32_t *pred) {
  for (int32_t v = 0; v < n; v++) pred[v] = -1;
  if (src < 0 || src >= n) return;
  pred[src] = src;
  for (int32_t u = 0; u < n; u++) {
    if (depth[u] < 0) continue;
    for (int32_t e = ro[u]; e < ro[u + 1]; e++) {
      const int32_t v = col[e];
      if (v != src && pred[v] == -1 && depth[v] == depth[u] + 1) pred[v] = u;
    }
  }
}

/* The reference's BFS validation (the five Graph500 rules listed at
 * Algorithms/BFS/bfs.cu:984-989, directed form of its edge rule at
 * bfs.cu:994-1012).  Returns 0 when (depth, pred) pass, else the first rule
 * broken:
 *   1 the root: depth[src] = 0, pred[src] = src, no other depth-0 vertex
 *     (with rule 2 this makes the tree acyclic: depth drops by one per step)
 *   2 each tree edge joins levels that differ by exactly one
 *   3 every edge u->v with u reached has v reached and depth[v] <= depth[u]+1
 *   4 the tree spans the reached set: pred >= 0 exactly where depth >= 0
 *   5 (pred[v], v) is an edge of the graph */
int32_t oracle_bfs_validate(int32_t n, const int32_t *ro, const int32_t *col,
                            int32_t src, const int32_t *depth,
                            const int32_t *pred) {
  if (src < 0 || src >= n || depth[src] != 0 || pred[src] != src) return 1;
  for (int32_t v = 0; v < n; v++)
    if (v != src && depth[v] == 0) return 1;
  for (int32_t v = 0; v < n; v++) {
    if (v == src || depth[v] < 0) continue;
    const int32_t p = pred[v];
    if (p < 0 || p >= n || depth[p] < 0 || depth[v] != depth[p] + 1) return 2;
  }
  for (int32_t u = 0; u < n; u++) {
    if (depth[u] < 0) continue;
    for (int32_t e = ro[u]; e < ro[u + 1]; e++) {
      const int32_t v = col[e];
      if (depth[v] < 0 || depth[v] > depth[u] + 1) return 3;
    }
  }
  for (int32_t v = 0; v < n; v++)
    if ((depth[v] >= 0) != (pred[v] >= 0)) return 4;
  unsigned char *found = (unsigned char *)calloc((size_t)(n > 0 ? n : 1), 1);
  for (int32_t u = 0; u < n; u++)
    for (int32_t e = ro[u]; e < ro[u + 1]; e++)
      if (pred[col[e]] == u) found[col[e]] = 1;
  int32_t rc = 0;
  for (int32_t v = 0; v < n && !rc; v++)
    if (v != src && depth[v] >= 0 && !found[v]) rc = 5;
  free(found);
  return rc;
}

/* Algorithms/SSSP/sssp.cu:126-183 -- selection Dijkstra exactly as written:
 * pick the unvisited vertex with the smallest dist < 100000000 (lowest id on
 * ties), relax its out-edges with 64-bit sums. */
void oracle_sssp_selection(int32_t n, const int32_t *ro, const int32_t *col,
                           const int32_t *w, int32_t src, int32_t *dist) {
  for (int32_t i = 0; i < n; i++) dist[i] = SSSP_INF;
  if (src < 0 || src >= n) return;
  unsigned char *visited = (unsigned char *)calloc((size_t)(n > 0 ? n : 1), 1);
  dist[src] = 0;
  for (;;) {
    int32_t u = -1, sd = SSSP_INF;
    for (int32_t i = 0; i < n; i++)
      if (!visited[i] && dist[i] < sd) {
        sd = dist[i];
        u = i;
      }
    if (u == -1) break;
    visited[u] = 1;
    for (int32_t e = ro[u]; e < ro[u + 1]; e++) {
      int32_t v = col[e];
      int64_t len = (int64_t)dist[u] + (int64_t)w[e];
      if (len < dist[v]) dist[v] = (int32_t)len;
    }
  }
  free(visited);
}

/* Binary min-heap of (dist, vertex) with lazy deletion. */
typedef struct {
  int64_t key;
  int32_t v;
} heap_item;

static void heap_push(heap_item *h, int64_t *size, heap_item it) {
  int64_t i = (*size)++;
  while (i > 0) {
    int64_t p = (i - 1) >> 1;
    if (h[p].key <= it.key) break;
    h[i] = h[p];
    i = p;
  }
  h[i] = it;
}

static heap_item heap_pop(heap_item *h, int64_t *size) {
  heap_item top = h[0];
  heap_item last = h[--(*size)];
  int64_t i = 0;
  for (;;) {
    int64_t c = 2 * i + 1;
    if (c >= *size) break;
    if (c + 1 < *size && h[c + 1].key < h[c].key) c++;
    if (last.key <= h[c].key) break;
    h[i] = h[c];
    i = c;
  }
  if (*size > 0) h[i] = last;
  return top;
}

/* Same fixed point as oracle_sssp_selection (unique shortest-path distances,
 * 100000000 for unreached, strict-< relaxation so nothing reaches the
 * sentinel), with a heap instead of the O(n) selection scan. */
void oracle_sssp_dijkstra(int32_t n, const int32_t *ro, const int32_t *col,
                          const int32_t *w, int32_t src, int32_t *dist) {
  for (int32_t i = 0; i < n; i++) dist[i] = SSSP_INF;
  if (src < 0 || src >= n) return;
  int64_t m = ro[n];
  heap_item *h = (heap_item *)malloc(sizeof(heap_item) * (size_t)(m + 2));
  unsigned char *done = (unsigned char *)calloc((size_t)n, 1);
  int64_t size = 0;
  dist[src] = 0;
  heap_item s = {0, src};
  heap_push(h, &size, s);
  while (size > 0) {
    heap_item it = heap_pop(h, &size);
    int32_t u = it.v;
    if (done[u] || it.key != dist[u]) continue;
    done[u] = 1;
    for (int32_t e = ro[u]; e < ro[u + 1]; e++) {
      int32_t v = col[e];
      int64_t len = (int64_t)dist[u] + (int64_t)w[e];
      if (len < dist[v]) {
        dist[v] = (int32_t)len;
        heap_item ni = {len, v};
        heap_push(h, &size, ni);
      }
    }
  }
  free(done);
  free(h);
}

/* Algorithms/CC/cc.cu:127-169 -- label = id, then in-place sweeps
 * label[v] = min(label[v], min_{nb} label[nb]) until no change. */
int32_t oracle_cc(int32_t n, const int32_t *ro, const int32_t *col,
                  int32_t *label) {
  for (int32_t i = 0; i < n; i++) label[i] = i;
  int32_t sweeps = 0;
  int changed = 1;
  while (changed) {
    changed = 0;
    sweeps++;
    for (int32_t v = 0; v < n; v++) {
      int32_t minnb = SSSP_INF;
      for (int32_t e = ro[v]; e < ro[v + 1]; e++) {
        int32_t l = label[col[e]];
        if (l < minnb) minnb = l;
      }
      if (minnb < label[v]) {
        label[v] = minnb;
        changed = 1;
      }
    }
  }
  return sweeps;
}

/* Algorithms/PageRank/PageRank.cu:108-161 -- in-place (Gauss-Seidel) float
 * sweeps over the CSC; 0.15 + 0.85*sum is formed in double then stored to
 * float, as the reference's mixed float/double expression does, and the
 * float |new - old| is compared with the double constant 0.01
 * (PageRank.cu:146). */
int32_t oracle_pr_reference(int32_t n, const int32_t *ro, const int32_t *co,
                            const int32_t *ri, float *rank) {
  for (int32_t i = 0; i < n; i++) rank[i] = 0.15f;
  int32_t iters = 0;
  int changed = 1;
  while (changed) {
    changed = 0;
    for (int32_t v = 0; v < n; v++) {
      float sum = 0.0f;
      for (int32_t e = co[v]; e < co[v + 1]; e++) {
        int32_t nb = ri[e];
        sum += rank[nb] / (float)(ro[nb + 1] - ro[nb]);
      }
      sum = (float)(0.15 + 0.85 * (double)sum);
      if ((double)fabsf(sum - rank[v]) >= 0.01) changed = 1;
      rank[v] = sum;
    }
    iters++;
  }
  return iters;
}

int32_t oracle_pagerank_f64(int32_t n, const int32_t *ro, const int32_t *col,
                            const double *p, const double *t, double d,
                            double tol, int32_t max_iter, double *r,
                            double *resid) {
  double *y = (double *)malloc(sizeof(double) * (size_t)n);
  double inv_n = 1.0 / (double)n;
  for (int32_t v = 0; v < n; v++) r[v] = t ? t[v] : inv_n;
  int32_t it = 0;
  double res = INFINITY;
  while (it < max_iter) {
    double dangling = 0.0;
    memset(y, 0, sizeof(double) * (size_t)n);
    for (int32_t u = 0; u < n; u++) {
      int32_t deg = ro[u + 1] - ro[u];
      if (deg == 0) {
        dangling += r[u];
        continue;
      }
      double inv_deg = 1.0 / (double)deg;
      for (int32_t e = ro[u]; e < ro[u + 1]; e++)
        y[col[e]] += (p ? p[e] : inv_deg) * r[u];
    }
    res = 0.0;
    for (int32_t v = 0; v < n; v++) {
      double tv = t ? t[v] : inv_n;
      double rn = d * (y[v] + dangling * inv_n) + (1.0 - d) * tv;
      res += fabs(rn - r[v]);
      r[v] = rn;
    }
    it++;
    if (res <= tol) break;
  }
  if (resid) *resid = res;
  free(y);
  return it;
}

int32_t oracle_pagerank_adjoint_f64(int32_t n, const int32_t *ro,
                                    const int32_t *col, const double *w,
                                    double d, const double *r, const double *c,
                                    double tol, int32_t max_iter,
                                    double *lambda, double *grad_w,
                                    double *grad_t) {
  double *ln = (double *)malloc(sizeof(double) * (size_t)n);
  double *S = (double *)malloc(sizeof(double) * (size_t)n);
  double inv_n = 1.0 / (double)n;
  for (int32_t u = 0; u < n; u++) {
    double s = 0.0;
    if (w)
      for (int32_t e = ro[u]; e < ro[u + 1]; e++) s += w[e];
    else
      s = (double)(ro[u + 1] - ro[u]);
    S[u] = s;
  }
  for (int32_t v = 0; v < n; v++) lambda[v] = c[v];
  int32_t it = 0;
  while (it < max_iter) {
    double total = 0.0;
    for (int32_t v = 0; v < n; v++) total += lambda[v];
    double res = 0.0;
    for (int32_t u = 0; u < n; u++) {
      double acc = 0.0;
      int32_t deg = ro[u + 1] - ro[u];
      if (deg == 0) {
        acc = total * inv_n;
      } else {
        for (int32_t e = ro[u]; e < ro[u + 1]; e++) {
          double puv = w ? w[e] / S[u] : 1.0 / S[u];
          acc += puv * lambda[col[e]];
        }
      }
      ln[u] = c[u] + d * acc;
      res += fabs(ln[u] - lambda[u]);
    }
    memcpy(lambda, ln, sizeof(double) * (size_t)n);
    it++;
    if (res <= tol) break;
  }
  if (grad_t)
    for (int32_t v = 0; v < n; v++) grad_t[v] = (1.0 - d) * lambda[v];
  if (grad_w) {
    for (int32_t u = 0; u < n; u++) {
      if (ro[u + 1] == ro[u]) continue;
      if (w) {
        double mean = 0.0;
        for (int32_t e = ro[u]; e < ro[u + 1]; e++)
          mean += (w[e] / S[u]) * lambda[col[e]];
        double scale = d * r[u] / S[u];
        for (int32_t e = ro[u]; e < ro[u + 1]; e++)
          grad_w[e] = scale * (lambda[col[e]] - mean);
      } else {
        for (int32_t e = ro[u]; e < ro[u + 1]; e++)
          grad_w[e] = d * r[u] * lambda[col[e]];
      }
    }
  }
  free(S);
  free(ln);
  return it;
}

/* ---- parallel CPU baselines (OpenMP; same outputs as the serial ones) ---- */
#ifdef _OPENMP
#include <omp.h>
#endif

int32_t oracle_threads(void) {
#ifdef _OPENMP
  return (int32_t)omp_get_max_threads();
#else
  return 1;
#endif
}

/* Direction-optimising level-synchronous BFS (top-down with a CAS on the
 * depth, bottom-up over unvisited vertices when the frontier's out-edges
 * exceed the unexplored edges / 14), output equal to oracle_bfs.  `in_ro`,
 * `in_col`: in-edges (the CSR itself for a symmetric graph). */
void oracle_bfs_omp(int32_t n, const int32_t *ro, const int32_t *col, const int32_t *in_ro, const int32_t *in_col,
                    int32_t src, int32_t *depth) {
  for (int32_t i = 0; i < n; i++) depth[i] = -1;
  if (src < 0 || src >= n) return;
  int32_t *front = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  int32_t *next = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  int64_t nf = 1, explored = 0;
  const int64_t m = ro[n];
  front[0] = src;
  depth[src] = 0;
  for (int32_t level = 0; nf > 0; level++) {
    int64_t mf = 0;
#pragma omp parallel for reduction(+ : mf) schedule(static)
    for (int64_t i = 0; i < nf; i++) mf += ro[front[i] + 1] - ro[front[i]];
    explored += mf;
    int64_t nn = 0;
    if (mf > (m - explored) / 14) {
      /* bottom-up: every unvisited vertex looks for a parent at this level */
#pragma omp parallel for reduction(+ : nn) schedule(dynamic, 1024)
      for (int32_t v = 0; v < n; v++) {
        if (depth[v] != -1) continue;
        for (int32_t k = in_ro[v]; k < in_ro[v + 1]; k++)
          if (depth[in_col[k]] == level) {
            depth[v] = -2 - level; /* found this level; published after the sweep */
            nn++;
            break;
          }
      }
      /* publish the level and compact the new frontier (same static chunks
         in both loops, so each thread's offset is its chunk's prefix) */
      const int32_t T = oracle_threads();
      int64_t *cnt = (int64_t *)calloc((size_t)T + 1, sizeof(int64_t));
#pragma omp parallel num_threads(T)
      {
#ifdef _OPENMP
        const int t = omp_get_thread_num();
#else
        const int t = 0;
#endif
        int64_t c = 0;
#pragma omp for schedule(static)
        for (int32_t v = 0; v < n; v++)
          if (depth[v] == -2 - level) c++;
        cnt[t + 1] = c;
#pragma omp barrier
#pragma omp single
        for (int32_t i = 0; i < T; i++) cnt[i + 1] += cnt[i];
        int64_t pos = cnt[t];
#pragma omp for schedule(static)
        for (int32_t v = 0; v < n; v++)
          if (depth[v] == -2 - level) {
            depth[v] = level + 1;
            next[pos++] = v;
          }
      }
      nn = cnt[T];
      free(cnt);
    } else {
#pragma omp parallel
      {
        int32_t *local = (int32_t *)malloc(sizeof(int32_t) * 4096);
        int32_t cnt = 0;
#pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < nf; i++) {
          const int32_t u = front[i];
          for (int32_t k = ro[u]; k < ro[u + 1]; k++) {
            const int32_t v = col[k];
            int32_t expect = -1;
            if (depth[v] == -1 &&
                __atomic_compare_exchange_n(&depth[v], &expect, level + 1, 0, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
              if (cnt == 4096) {
                int64_t at = __atomic_fetch_add(&nn, (int64_t)cnt, __ATOMIC_RELAXED);
                memcpy(next + at, local, sizeof(int32_t) * (size_t)cnt);
                cnt = 0;
              }
              local[cnt++] = v;
            }
          }
        }
        int64_t at = __atomic_fetch_add(&nn, (int64_t)cnt, __ATOMIC_RELAXED);
        memcpy(next + at, local, sizeof(int32_t) * (size_t)cnt);
        free(local);
      }
    }
    int32_t *t = front;
    front = next;
    next = t;
    nf = nn;
  }
  free(front);
  free(next);
}

/* Bucketed SSSP (near-far: relax from the vertices below the current bound
 * until none is left, then move the bound to the bucket of the smallest
 * pending distance), CAS-min on int32 distances, 64-bit sums and the
 * 100000000 sentinel as oracle_sssp_selection; output equal to it. */
void oracle_sssp_omp(int32_t n, const int32_t *ro, const int32_t *col, const int32_t *w, int32_t src, int32_t delta,
                     int32_t *dist) {
  for (int32_t i = 0; i < n; i++) dist[i] = SSSP_INF;
  if (src < 0 || src >= n) return;
  unsigned char *dirty = (unsigned char *)calloc((size_t)(n > 0 ? n : 1), 1);
  int32_t *front = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  dist[src] = 0;
  dirty[src] = 1;
  int64_t bound = delta;
  for (;;) {
    int64_t nf = 0, remain = 0;
    int32_t dmin = INT32_MAX;
#pragma omp parallel
    {
      int32_t *local = (int32_t *)malloc(sizeof(int32_t) * 4096);
      int32_t cnt = 0, lmin = INT32_MAX;
      int64_t lrem = 0;
#pragma omp for schedule(static)
      for (int32_t v = 0; v < n; v++) {
        if (!dirty[v]) continue;
        if (dist[v] < bound) {
          dirty[v] = 0;
          if (cnt == 4096) {
            int64_t at = __atomic_fetch_add(&nf, (int64_t)cnt, __ATOMIC_RELAXED);
            memcpy(front + at, local, sizeof(int32_t) * (size_t)cnt);
            cnt = 0;
          }
          local[cnt++] = v;
        } else {
          lrem++;
          if (dist[v] < lmin) lmin = dist[v];
        }
      }
      int64_t at = __atomic_fetch_add(&nf, (int64_t)cnt, __ATOMIC_RELAXED);
      memcpy(front + at, local, sizeof(int32_t) * (size_t)cnt);
      __atomic_fetch_add(&remain, lrem, __ATOMIC_RELAXED);
#pragma omp critical
      if (lmin < dmin) dmin = lmin;
      free(local);
    }
    if (nf == 0) {
      if (remain == 0) break;
      bound = ((int64_t)dmin / delta + 1) * delta;
      continue;
    }
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < nf; i++) {
      const int32_t u = front[i];
      const int64_t du = __atomic_load_n(&dist[u], __ATOMIC_RELAXED);
      for (int32_t k = ro[u]; k < ro[u + 1]; k++) {
        const int32_t v = col[k];
        const int64_t nd = du + w[k];
        if (nd >= SSSP_INF) continue;
        int32_t cur = __atomic_load_n(&dist[v], __ATOMIC_RELAXED);
        while (nd < cur) {
          if (__atomic_compare_exchange_n(&dist[v], &cur, (int32_t)nd, 0, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
            dirty[v] = 1;
            break;
          }
        }
      }
    }
  }
  free(dirty);
  free(front);
}

/* ---- finite differences of the FP64 PageRank loss (see oracle.h) ---- */
#define FD_K 64 /* perturbations per sweep: 32 edges x (+h, -h) */

int32_t oracle_pagerank_fd_f64(int32_t n, const int32_t *ro, const int32_t *col,
                               const double *w, double d, const double *r,
                               const double *c, int32_t ne, const int64_t *edges,
                               const double *h, double tol_rel, int32_t max_iter,
                               double *fd) {
  const int64_t m = ro[n];
  const double inv_n = 1.0 / (double)n;
  /* In-edges with their transition weights P_yx (CSC by counting sort). */
  int32_t *co = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
  int32_t *src = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
  double *pw = (double *)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
  double *S = (double *)malloc(sizeof(double) * (size_t)n);
  int32_t *fill = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
  unsigned char *dang = (unsigned char *)malloc((size_t)n);
  for (int32_t u = 0; u < n; u++) {
    double s = 0.0;
    for (int32_t e = ro[u]; e < ro[u + 1]; e++) s += w[e];
    S[u] = s;
    dang[u] = ro[u + 1] == ro[u];
  }
  for (int64_t e = 0; e < m; e++) co[col[e] + 1]++;
  for (int32_t v = 0; v < n; v++) co[v + 1] += co[v];
  memcpy(fill, co, sizeof(int32_t) * (size_t)n);
  for (int32_t u = 0; u < n; u++)
    for (int32_t e = ro[u]; e < ro[u + 1]; e++) {
      int32_t k = fill[col[e]]++;
      src[k] = u;
      pw[k] = w[e] / S[u];
    }
  double *D = (double *)malloc(sizeof(double) * (size_t)n * FD_K);
  double *Dn = (double *)malloc(sizeof(double) * (size_t)n * FD_K);
  int32_t max_sweeps = 0;
  for (int32_t base = 0; base < ne; base += FD_K / 2) {
    const int32_t kb = (ne - base < FD_K / 2 ? ne - base : FD_K / 2) * 2; /* active perturbations */
    int32_t pu[FD_K];
    double sh[FD_K], scale[FD_K]; /* signed h; s h / (S_u + s h) */
    int64_t pe[FD_K];
    for (int32_t j = 0; j < FD_K; j++) {
      const int32_t i = base + j / 2;
      if (j >= kb) {
        pu[j] = -1;
        continue;
      }
      const int64_t e = edges[i];
      int32_t lo = 0, hi = n; /* row of edge e */
      while (hi - lo > 1) {
        int32_t mid = (lo + hi) / 2;
        if (ro[mid] <= e) lo = mid; else hi = mid;
      }
      pu[j] = lo;
      pe[j] = e;
      sh[j] = (j & 1) ? -h[i] : h[i];
      scale[j] = sh[j] / (S[lo] + sh[j]);
    }
    /* P'_u. - P_u. = scale * (e_v - P_u.) on row u (v = the edge's target). */
    memset(D, 0, sizeof(double) * (size_t)n * FD_K);
    double diff[FD_K], norm[FD_K];
    int32_t sweep = 0;
    for (;;) {
      double dsum[FD_K], du[FD_K];
      for (int32_t j = 0; j < FD_K; j++) {
        dsum[j] = 0.0;
        du[j] = pu[j] >= 0 ? D[(size_t)pu[j] * FD_K + j] : 0.0;
      }
      for (int32_t y = 0; y < n; y++)
        if (dang[y])
          for (int32_t j = 0; j < FD_K; j++) dsum[j] += D[(size_t)y * FD_K + j];
#pragma omp parallel for schedule(dynamic, 2048)
      for (int32_t x = 0; x < n; x++) {
        double acc[FD_K];
        for (int32_t j = 0; j < FD_K; j++) acc[j] = dsum[j] * inv_n;
        for (int32_t k = co[x]; k < co[x + 1]; k++) {
          const double p = pw[k];
          const double *dy = D + (size_t)src[k] * FD_K;
#pragma omp simd
          for (int32_t j = 0; j < FD_K; j++) acc[j] += p * dy[j];
        }
        double *o = Dn + (size_t)x * FD_K;
        for (int32_t j = 0; j < FD_K; j++) o[j] = d * acc[j];
      }
      /* The changed row: d (P'_u. - P_u.) (D_u + r_u), applied on row u's targets. */
      for (int32_t j = 0; j < kb; j++) {
        const int32_t u = pu[j];
        const double f = d * scale[j] * (du[j] + r[u]);
        for (int32_t e = ro[u]; e < ro[u + 1]; e++) Dn[(size_t)col[e] * FD_K + j] -= f * (w[e] / S[u]);
        Dn[(size_t)col[pe[j]] * FD_K + j] += f;
      }
      for (int32_t j = 0; j < FD_K; j++) diff[j] = norm[j] = 0.0;
#pragma omp parallel
      {
        double ld[FD_K], ln[FD_K];
        for (int32_t j = 0; j < FD_K; j++) ld[j] = ln[j] = 0.0;
#pragma omp for schedule(static)
        for (int32_t x = 0; x < n; x++)
          for (int32_t j = 0; j < FD_K; j++) {
            const double a = Dn[(size_t)x * FD_K + j];
            ld[j] += fabs(a - D[(size_t)x * FD_K + j]);
            ln[j] += fabs(a);
          }
#pragma omp critical
        for (int32_t j = 0; j < FD_K; j++) {
          diff[j] += ld[j];
          norm[j] += ln[j];
        }
      }
      double *t = D;
      D = Dn;
      Dn = t;
      sweep++;
      int conv = 1;
      for (int32_t j = 0; j < kb; j++)
        if (diff[j] > tol_rel * norm[j]) conv = 0;
      if (conv || sweep >= max_iter) break;
    }
    if (sweep > max_sweeps) max_sweeps = sweep;
    for (int32_t j = 0; j < kb; j += 2) {
      double lp = 0.0, lm = 0.0;
      for (int32_t x = 0; x < n; x++) {
        lp += c[x] * D[(size_t)x * FD_K + j];
        lm += c[x] * D[(size_t)x * FD_K + j + 1];
      }
      fd[base + j / 2] = (lp - lm) / (2.0 * sh[j]);
    }
  }
  free(D);
  free(Dn);
  free(dang);
  free(fill);
  free(S);
  free(pw);
  free(src);
  free(co);
  return max_sweeps;
}

int32_t oracle_cc_omp(int32_t n, const int32_t *ro, const int32_t *col,
                      int32_t *label) {
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < n; i++) label[i] = i;
  int32_t sweeps = 0;
  int changed = 1;
  while (changed) {
    changed = 0;
    sweeps++;
#pragma omp parallel for schedule(static) reduction(| : changed)
    for (int32_t v = 0; v < n; v++) {
      int32_t minnb = __atomic_load_n(&label[v], __ATOMIC_RELAXED);
      const int32_t own = minnb;
      for (int32_t e = ro[v]; e < ro[v + 1]; e++) {
        const int32_t l = __atomic_load_n(&label[col[e]], __ATOMIC_RELAXED);
        if (l < minnb) minnb = l;
      }
      if (minnb < own) {
        __atomic_store_n(&label[v], minnb, __ATOMIC_RELAXED);
        changed = 1;
      }
    }
  }
  return sweeps;
}

int32_t oracle_pagerank_f64_omp(int32_t n, const int32_t *ro, const int32_t *co,
                                const int32_t *ri, const double *p_csc,
                                const double *t, double d, double tol,
                                int32_t max_iter, double *r, double *resid) {
  double *x = (double *)malloc(sizeof(double) * (size_t)n); /* r_u * P_u. factor per source */
  double *y = (double *)malloc(sizeof(double) * (size_t)n);
  const double inv_n = 1.0 / (double)n;
#pragma omp parallel for schedule(static)
  for (int32_t v = 0; v < n; v++) r[v] = t ? t[v] : inv_n;
  int32_t it = 0;
  double res = INFINITY;
  while (it < max_iter) {
    double dangling = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : dangling)
    for (int32_t u = 0; u < n; u++) {
      const int32_t deg = ro[u + 1] - ro[u];
      if (deg == 0) dangling += r[u];
      x[u] = p_csc ? r[u] : (deg ? r[u] / (double)deg : 0.0);
    }
    res = 0.0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : res)
    for (int32_t v = 0; v < n; v++) {
      double acc = 0.0;
      if (p_csc)
        for (int32_t k = co[v]; k < co[v + 1]; k++) acc += p_csc[k] * x[ri[k]];
      else
        for (int32_t k = co[v]; k < co[v + 1]; k++) acc += x[ri[k]];
      const double tv = t ? t[v] : inv_n;
      y[v] = d * (acc + dangling * inv_n) + (1.0 - d) * tv;
      res += fabs(y[v] - r[v]);
    }
    memcpy(r, y, sizeof(double) * (size_t)n);
    it++;
    if (res <= tol) break;
  }
  if (resid) *resid = res;
  free(x);
  free(y);
  return it;
}
