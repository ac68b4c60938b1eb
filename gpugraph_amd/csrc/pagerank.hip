// PageRank forward (FP32 compensated gather, FP64 dangling mass and L1
// residual) and its adjoint, on gfx950.
//
// Forward:  r = d (P^T r + D/n 1) + (1-d) t,  D = sum of r over dangling
//           vertices, P_uv = w_uv / S_u (F32 values) or 1/outdeg(u).
//           One sweep = row sums over the CSC in-edge lists
//           (GASengine::seg_rows_launch: merge-path tiles, Neumaier-
//           compensated FP32, fixed fold order), writing r_next and, unweighted, the contribution
//           x = r_next / outdeg used by the next sweep; the same pass adds
//           |r_next - r| and the dangling mass of r_next in FP64 (one atomic
//           per workgroup).  A one-thread finalize kernel closes the
//           iteration on the device (residual, D, iteration count, done
//           flag); sweeps early-exit once done, so the host checks the state
//           only every kItersPerCheck iterations.
// Adjoint:  for L = c^T r,  lambda = c + d (P lambda + delta (1^T lambda)/n)
//           (the transposed fixed point: row reductions over CSR out-edges),
//           dL/dt = (1-d) lambda, and for raw weights
//           dL/dw_uv = d r_u / S_u (lambda_v - sum_v' P_uv' lambda_v')
//           (unweighted graphs: dL/dP_uv = d r_u lambda_v).  The forward
//           ranks are an input (stored by the caller), not recomputed.
//
// Reference: the pagerank program's gather/apply (Algorithms/PageRank/
// PageRank.h:197-281) computes 0.15 + 0.85 sum r_u/outdeg_u without a
// dangling term or teleport vector; that form runs through the generic GAS
// path (gg_pagerank_reference).  The adjoint has no reference counterpart.
//
// Byte model per sweep: unweighted 20n + 8m (+4n with a teleport vector):
// per vertex offsets, old/new rank, new contribution, 1/outdeg; per edge
// the source id and its contribution.  Weighted 16n + 12m (+4n): per edge
// source id, P_uv and r_u.  Adjoint sweep 24n + 8m (weighted + 4m): per
// vertex offsets, old/new lambda, c, row sum, 1/outdeg; per edge target id,
// its lambda (and P_uv).  Gradient pass 16m + 16n.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstring>


#include "gasengine/rowreduce.h"
#include "gasengine/segreduce.h"
#include "gasengine/xcd_slices.h"
#include "internal.h"

namespace gg {
namespace {

using namespace GASengine;
constexpr int kItersPerCheck = 8;

constexpr int kPrStride = 16;  // doubles between the stripes of a sum (128 bytes)

struct PrState {
  double aux_cur;             // dangling mass (forward) / sum of lambda (adjoint) of the current vector
  double scale;               // residual scale: 1 (forward) or ||c||_1 (adjoint)
  double resid_last;
  int32_t iters, done;
  // Per-workgroup FP64 sums, striped (see GASengine::block_counter_add):
  // accumulated by a sweep (or k_pr_init), folded by k_pr_finalize /
  // k_pr_start.
  double resid_acc[kStripes * kPrStride], aux_acc[kStripes * kPrStride], scale_acc[kStripes * kPrStride];
};

// Fold a striped FP64 sum and clear its stripes (one thread).
__device__ __forceinline__ double take_stripes(double *acc) {
  double t = 0.0;
  for (int q = 0; q < kStripes; q++) {
    t += acc[q * kPrStride];
    acc[q * kPrStride] = 0.0;
  }
  return t;
}

template <int BLOCK>
__device__ __forceinline__ void block_add_f64(double *p, double v) {
  __shared__ double red[BLOCK / kWave];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if (lane_id() == 0) red[threadIdx.x / kWave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int w = 0; w < BLOCK / kWave; w++) t += red[w];
    if (t != 0.0) atomicAdd(stripe_of<kPrStride>(p), t);
  }
  __syncthreads();
}

struct PrForwardOp {
  const int32_t *off;   // CSC column offsets
  const int32_t *src;   // CSC row indices (in-neighbours)
  const float *pcsc;    // P_uv per CSC slot (weighted) or nullptr
  const float *inv_out; // 1/outdeg (0: dangling)
  float *r[2];
  float *x[2];          // contributions r/outdeg (unweighted)
  const float *tele;    // teleport vector or nullptr (uniform)
  float d, uniform_tele;
  double inv_n;
  PrState *st;
  // This sweep's vectors, picked in begin() by the iteration's parity with
  // constant indices (a run-time index into r / x would keep the whole op in
  // scratch memory).
  const float *rc, *xc;
  float *rn, *xn;
  float dn;
  double resid, dang;

  __device__ __forceinline__ void begin() {
    const bool odd = st->iters & 1;
    rc = odd ? r[1] : r[0];
    rn = odd ? r[0] : r[1];
    xc = odd ? x[1] : x[0];
    xn = odd ? x[0] : x[1];
    dn = (float)(st->aux_cur * inv_n);
    resid = 0.0;
    dang = 0.0;
  }
  __device__ __forceinline__ int32_t row_begin(int32_t v) { return off[v]; }
  __device__ __forceinline__ int32_t row_end(int32_t v) { return off[v + 1]; }
  // Staged form (segreduce.h): source and P_uv load with the tile, the
  // contribution r_u P_uv (or x_u) is gathered after; r_v and 1/outdeg_v
  // (read-only during a sweep) are prefetched for item_done.
  struct Edge {
    int32_t u;
    float p;
  };
  struct State {
    float r_old, io;
  };
  // The edge streams are read once per sweep: non-temporal loads keep them from
  // displacing the gathered contributions in L2.
  __device__ __forceinline__ Edge load_edge(int32_t k) {
    return {__builtin_nontemporal_load(&src[k]), pcsc ? __builtin_nontemporal_load(&pcsc[k]) : 0.0f};
  }
  __device__ __forceinline__ float gather(const Edge &e) { return pcsc ? rc[e.u] * e.p : xc[e.u]; }
  __device__ __forceinline__ State vertex_state(int32_t v) { return {rc[v], inv_out[v]}; }
  __device__ __forceinline__ void item_done(int32_t v, float sum, const State &st) {
    const float tv = tele ? (1.0f - d) * tele[v] : uniform_tele;
    const float rv = d * (sum + dn) + tv;
    rn[v] = rv;
    if (!pcsc) xn[v] = rv * st.io;
    resid += fabs((double)rv - (double)st.r_old);
    if (st.io == 0.0f) dang += rv;
  }
  // Per-edge form (XCD-sliced sweeps).
  __device__ __forceinline__ float edge(int32_t, int32_t k) { return gather(load_edge(k)); }
  __device__ __forceinline__ void item_done(int32_t v, float sum) { item_done(v, sum, vertex_state(v)); }
  __device__ __forceinline__ void finish() {
    block_add_f64<kRowBlock>(st->resid_acc, resid);
    block_add_f64<kRowBlock>(st->aux_acc, dang);
  }
};

struct PrAdjointOp {
  const int32_t *off;   // CSR row offsets
  const int32_t *dst;   // CSR column indices
  const float *pcsr;    // P_uv per CSR slot (weighted) or nullptr
  const float *inv_out;
  float *lam[2];
  const float *c;
  float *rowsum;        // sum_v P_uv lambda_v of the vector read by this sweep
  float d;
  double inv_n;
  PrState *st;
  const float *lc;  // this sweep's lambda and the next one (picked in begin())
  float *lnx;
  float ln;
  double resid, lsum;

  __device__ __forceinline__ void begin() {
    const bool odd = st->iters & 1;
    lc = odd ? lam[1] : lam[0];
    lnx = odd ? lam[0] : lam[1];
    ln = (float)(st->aux_cur * inv_n);
    resid = 0.0;
    lsum = 0.0;
  }
  __device__ __forceinline__ int32_t row_begin(int32_t u) { return off[u]; }
  __device__ __forceinline__ int32_t row_end(int32_t u) { return off[u + 1]; }
  // Staged form (segreduce.h): target and P_uv load with the tile, lambda_v
  // is gathered after; lambda_u, 1/outdeg_u and c_u (read-only during a
  // sweep) are prefetched for item_done.
  struct Edge {
    int32_t v;
    float p;
  };
  struct State {
    float l_old, io, cu;
  };
  __device__ __forceinline__ Edge load_edge(int32_t e) {
    return {__builtin_nontemporal_load(&dst[e]), pcsr ? __builtin_nontemporal_load(&pcsr[e]) : 0.0f};
  }
  __device__ __forceinline__ float gather(const Edge &e) {
    const float l = lc[e.v];
    return pcsr ? e.p * l : l;
  }
  __device__ __forceinline__ State vertex_state(int32_t u) { return {lc[u], inv_out[u], c[u]}; }
  __device__ __forceinline__ void item_done(int32_t u, float sum, const State &st) {
    const float su = pcsr ? sum : sum * st.io;
    const float ln_u = st.cu + d * (su + (st.io == 0.0f ? ln : 0.0f));
    lnx[u] = ln_u;
    rowsum[u] = su;
    resid += fabs((double)ln_u - (double)st.l_old);
    lsum += ln_u;
  }
  // Per-edge form (XCD-sliced sweeps).
  __device__ __forceinline__ float edge(int32_t, int32_t e) { return gather(load_edge(e)); }
  __device__ __forceinline__ void item_done(int32_t u, float sum) { item_done(u, sum, vertex_state(u)); }
  __device__ __forceinline__ void finish() {
    block_add_f64<kRowBlock>(st->resid_acc, resid);
    block_add_f64<kRowBlock>(st->aux_acc, lsum);
  }
};

__global__ void k_pr_finalize(PrState *st, double tol, int32_t max_iter) {
  if (st->done) return;
  st->resid_last = take_stripes(st->resid_acc);
  st->aux_cur = take_stripes(st->aux_acc);
  st->iters++;
  if (st->resid_last <= tol * st->scale || st->iters >= max_iter) st->done = 1;
}

// Start vector (forward: teleport or uniform; adjoint: c) and its aux sum.
__global__ void __launch_bounds__(kRowBlock)
k_pr_init(float *r0, float *x0, const float *tele, const float *inv_out, int32_t n, PrState *st, int adjoint,
          const float *c) {
  double aux = 0.0, mag = 0.0;
  for (int64_t v = (int64_t)blockIdx.x * kRowBlock + threadIdx.x; v < n; v += (int64_t)gridDim.x * kRowBlock) {
    if (adjoint) {
      const float l = c[v];
      r0[v] = l;
      aux += l;
      mag += fabs((double)l);
    } else {
      const float rv = tele ? tele[v] : (float)(1.0 / (double)n);
      r0[v] = rv;
      if (x0) x0[v] = rv * inv_out[v];
      if (inv_out[v] == 0.0f) aux += rv;
    }
  }
  block_add_f64<kRowBlock>(st->aux_acc, aux);
  block_add_f64<kRowBlock>(st->scale_acc, mag);
}

__global__ void k_pr_start(PrState *st, int adjoint) {
  st->scale = take_stripes(st->scale_acc);
  if (!adjoint || st->scale == 0.0) st->scale = 1.0;
  st->aux_cur = take_stripes(st->aux_acc);
  (void)take_stripes(st->resid_acc);
  st->resid_last = INFINITY;
  st->iters = 0;
  st->done = 0;
}

__global__ void k_inv_outdeg(const int32_t *__restrict__ ro, float *__restrict__ inv, int32_t n) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
    const int32_t dgr = ro[v + 1] - ro[v];
    inv[v] = dgr > 0 ? (float)(1.0 / (double)dgr) : 0.0f;
  }
}

// S_u = sum of u's out-edge weights (FP64, in a fixed order), then
// P_uv = w_uv / S_u rounded to FP32 for every CSR slot of the row.  Rows come
// from the CSR degree bins: MODE 0 one thread per row (small bin), 1 one
// wave (mid), 2 one workgroup (large), so a hub row is not one thread's job.
template <int MODE>
__global__ void __launch_bounds__(kRowBlock)
k_row_normalise(const int32_t *__restrict__ ro, const float *__restrict__ w, const int32_t *__restrict__ list,
                int32_t count, double *__restrict__ S, float *__restrict__ pcsr) {
  __shared__ double red[kRowBlock / kWave];
  constexpr int step = MODE == 0 ? 1 : MODE == 1 ? kWave : kRowBlock;
  const int me = MODE == 0 ? 0 : MODE == 1 ? lane_id() : (int)threadIdx.x;
  int64_t i, stride;
  if (MODE == 0) {
    i = (int64_t)blockIdx.x * kRowBlock + threadIdx.x;
    stride = (int64_t)gridDim.x * kRowBlock;
  } else if (MODE == 1) {
    i = (int64_t)blockIdx.x * (kRowBlock / kWave) + threadIdx.x / kWave;
    stride = (int64_t)gridDim.x * (kRowBlock / kWave);
  } else {
    i = blockIdx.x;
    stride = gridDim.x;
  }
  for (; i < count; i += stride) {
    const int32_t u = list[i];
    const int32_t b = ro[u], e = ro[u + 1];
    double acc = 0.0;
    for (int32_t k = b + me; k < e; k += step) acc += (double)w[k];
    if (MODE >= 1) {
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
      acc = __shfl(acc, 0);  // every lane uses lane 0's rounding
    }
    if (MODE == 2) {
      if (lane_id() == 0) red[threadIdx.x / kWave] = acc;
      __syncthreads();
      acc = 0.0;
      for (int q = 0; q < kRowBlock / kWave; q++) acc += red[q];
      __syncthreads();
    }
    if (me == 0) S[u] = acc;
    for (int32_t k = b + me; k < e; k += step) pcsr[k] = (float)((double)w[k] / acc);
  }
}

// P per CSC slot k (source u = src[k]) from the CSC-ordered raw weights.
__global__ void k_csc_normalise(const int32_t *__restrict__ src, const float *__restrict__ wc,
                                const double *__restrict__ S, float *__restrict__ pcsc, int64_t m) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += (int64_t)gridDim.x * blockDim.x)
    pcsc[k] = (float)((double)wc[k] / S[src[k]]);
}

// Gradient pass per row u: s_u = sum_e P_e lambda_v with the final lambda,
// then grad[e] = scale_u * (lambda_v - s_u) (weighted, scale_u = d r_u / S_u)
// or d r_u lambda_v (unweighted).  One thread per row; rows are short on
// average and the pass runs once per backward call.
__global__ void __launch_bounds__(kRowBlock)
k_pr_grad_rows(const int32_t *__restrict__ ro, const int32_t *__restrict__ col, const float *__restrict__ pcsr,
               const double *__restrict__ S, const float *__restrict__ rank, const float *__restrict__ lam, float d,
               const int32_t *__restrict__ list, int32_t count, float *__restrict__ grad) {
  for (int64_t i = (int64_t)blockIdx.x * kRowBlock + threadIdx.x; i < count; i += (int64_t)gridDim.x * kRowBlock) {
    const int32_t u = list[i];
    const int32_t b = ro[u], e = ro[u + 1];
    if (pcsr) {
      KSum acc;
      for (int32_t k = b; k < e; k++) acc.add(pcsr[k] * lam[col[k]]);
      const float su = acc.value();
      const float scale = (float)((double)d * (double)rank[u] / S[u]);
      for (int32_t k = b; k < e; k++) grad[k] = scale * (lam[col[k]] - su);
    } else {
      const float scale = d * rank[u];
      for (int32_t k = b; k < e; k++) grad[k] = scale * lam[col[k]];
    }
  }
}

// Same for the mid / large rows: one wave per row, lanes stride the row.
__global__ void __launch_bounds__(kRowBlock)
k_pr_grad_rows_wave(const int32_t *__restrict__ ro, const int32_t *__restrict__ col, const float *__restrict__ pcsr,
                    const double *__restrict__ S, const float *__restrict__ rank, const float *__restrict__ lam, float d,
                    const int32_t *__restrict__ list, int32_t count, float *__restrict__ grad) {
  const int64_t waves = (int64_t)gridDim.x * (kRowBlock / kWave);
  for (int64_t i = (int64_t)blockIdx.x * (kRowBlock / kWave) + threadIdx.x / kWave; i < count; i += waves) {
    const int32_t u = list[i];
    const int32_t b = ro[u], e = ro[u + 1];
    if (pcsr) {
      KSum acc;
      for (int32_t k = b + lane_id(); k < e; k += kWave) acc.add(pcsr[k] * lam[col[k]]);
      wave_merge(acc);
      const float su = __shfl(acc.value(), 0);
      const float scale = (float)((double)d * (double)rank[u] / S[u]);
      for (int32_t k = b + lane_id(); k < e; k += kWave) grad[k] = scale * (lam[col[k]] - su);
    } else {
      const float scale = d * rank[u];
      for (int32_t k = b + lane_id(); k < e; k += kWave) grad[k] = scale * lam[col[k]];
    }
  }
}

__global__ void k_scale(const float *__restrict__ in, float a, float *__restrict__ out, int32_t n) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x)
    out[v] = a * in[v];
}

}  // namespace

// Degree bins over CSC (dir 0) or CSR (dir 1), cached per graph.
int ensure_bins(gg_graph *g, int dir) {
  if (g->bins_ready[dir]) return GG_OK;
  const int32_t *off = dir == 0 ? g->dg.column_offsets : g->dg.row_offsets;
  for (int b = 0; b < 3; b++) GG_TRY(g->bins[dir][b].ensure(4 * (size_t)g->n));
  GG_TRY(g->counters.ensure(kCounterWords * sizeof(unsigned long long)));
  int32_t *cnt = reinterpret_cast<int32_t *>(g->counters.as<unsigned long long>() + 48);
  GG_HIP(fill_async<int32_t>(cnt, 0, 3, g->stream));
  hipLaunchKernelGGL(k_build_bins<0>, dim3(grid_for(g->n, kRowBlock, 4096)), dim3(kRowBlock), 0, g->stream, off, g->n,
                     g->bins[dir][0].as<int32_t>(), g->bins[dir][1].as<int32_t>(), g->bins[dir][2].as<int32_t>(), cnt);
  GG_HIP(hipGetLastError());
  GG_HIP(hipMemcpyAsync(g->bin_count[dir], cnt, 3 * sizeof(int32_t), hipMemcpyDeviceToHost, g->stream));
  GG_HIP(hipStreamSynchronize(g->stream));
  g->bins_ready[dir] = true;
  return GG_OK;
}

// Transition weights: pr[0] = 1/outdeg (float), and for F32-valued graphs
// pr[1] = S (double), pr[2] = P per CSR slot, pr[3] = P per CSC slot.
int ensure_transition(gg_graph *g) {
  const bool weighted = g->value_dtype == GG_VALUES_F32;
  if (g->pr_weights_version == g->values_version && g->pr[0].p) return GG_OK;
  const int32_t n = g->n;
  const int64_t m = g->m;
  hipStream_t s = g->stream;
  GG_TRY(g->pr[0].ensure(4 * (size_t)n));
  hipLaunchKernelGGL(k_inv_outdeg, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, g->dg.row_offsets,
                     g->pr[0].as<float>(), n);
  GG_HIP(hipGetLastError());
  if (weighted) {
    GG_TRY(g->pr[1].ensure(8 * (size_t)n));
    GG_TRY(g->pr[2].ensure(4 * (size_t)std::max<int64_t>(m, 1)));
    GG_TRY(g->pr[3].ensure(4 * (size_t)std::max<int64_t>(m, 1)));
    GG_TRY(ensure_bins(g, 1));
    const float *w = g->val.as<float>();
    double *S = g->pr[1].as<double>();
    float *pcsr = g->pr[2].as<float>();
    const int32_t *ro = g->dg.row_offsets;
    if (const int32_t c0 = g->bin_count[1][0])
      hipLaunchKernelGGL(k_row_normalise<0>, dim3(grid_for(c0, kRowBlock, 8192)), dim3(kRowBlock), 0, s, ro, w,
                         g->bins[1][0].as<int32_t>(), c0, S, pcsr);
    if (const int32_t c1 = g->bin_count[1][1])
      hipLaunchKernelGGL(k_row_normalise<1>, dim3(grid_for((int64_t)c1 * kWave, kRowBlock, 8192)), dim3(kRowBlock), 0,
                         s, ro, w, g->bins[1][1].as<int32_t>(), c1, S, pcsr);
    if (const int32_t c2 = g->bin_count[1][2])
      hipLaunchKernelGGL(k_row_normalise<2>, dim3(c2 < 4096 ? c2 : 4096), dim3(kRowBlock), 0, s, ro, w,
                         g->bins[1][2].as<int32_t>(), c2, S, pcsr);
    GG_HIP(hipGetLastError());
    const void *wc = nullptr;
    GG_TRY(csc_values(g, &wc));
    hipLaunchKernelGGL(k_csc_normalise, dim3(grid_for(m, 256, 16384)), dim3(256), 0, s, g->dg.row_indices,
                       static_cast<const float *>(wc), g->pr[1].as<double>(), g->pr[3].as<float>(), m);
    GG_HIP(hipGetLastError());
  }
  GG_HIP(hipStreamSynchronize(s));
  g->pr_weights_version = g->values_version;
  return GG_OK;
}

// Whether PageRank sweeps on this graph use the XCD-sliced rows.  Auto (-1)
// means off: on R-MAT scale 22 (average in-degree 16) the eight sub-rows per
// vertex cost more in offsets and partials than the L2-resident gathers save
// (19.3 ms vs 12.9 ms per forward solve against the merge-path sweep);
// gg_autotune tries both.
bool use_slices(const gg_graph *g) { return g->opt.xcd_slices == 1; }

// Slice structure of direction dir (0: CSC in-edges, 1: CSR out-edges),
// built once per graph; *L describes it for the sweep kernels.
int ensure_slices(gg_graph *g, int dir, SliceLists *L) {
  const int32_t n = g->n;
  const int64_t m = g->m;
  hipStream_t s = g->stream;
  const int32_t *off = dir == 0 ? g->dg.column_offsets : g->dg.row_offsets;
  const int32_t *nbr = dir == 0 ? g->dg.row_indices : g->dg.column_indices;
  const int64_t cells = (int64_t)kXcdSlices * n;
  if (!g->sl_ready[dir]) {
    DevBuf cnt, tmp;
    GG_TRY(cnt.ensure(4 * (size_t)(cells + 1)));
    GG_TRY(g->sl_off[dir].ensure(4 * (size_t)(cells + 1)));
    GG_TRY(g->sl_idx[dir].ensure(4 * (size_t)std::max<int64_t>(m, 1)));
    GG_TRY(g->sl_perm[dir].ensure(4 * (size_t)std::max<int64_t>(m, 1)));
    GG_HIP(fill_async<int32_t>(cnt.as<int32_t>() + cells, 0, 1, s));
    hipLaunchKernelGGL(k_slice_count<0>, dim3(grid_for((int64_t)n * kWave, 256, 16384)), dim3(256), 0, s, off, nbr, n,
                       cnt.as<int32_t>());
    GG_HIP(hipGetLastError());
    GG_TRY(tmp.ensure(4 * (size_t)scan_totals_entries(cells + 1)));
    GG_HIP(exclusive_scan_i32(cnt.as<int32_t>(), g->sl_off[dir].as<int32_t>(), cells + 1, tmp.as<int32_t>(), s));
    hipLaunchKernelGGL(k_slice_fill<0>, dim3(grid_for((int64_t)n * kWave, 256, 16384)), dim3(256), 0, s, off, nbr, n,
                       g->sl_off[dir].as<int32_t>(), g->sl_idx[dir].as<int32_t>(), g->sl_perm[dir].as<int32_t>());
    GG_HIP(hipGetLastError());
    // Long sub-rows: count, lay out per slice, fill.
    int32_t *ctr = cnt.as<int32_t>();  // reused: 16 counters + 18 bases
    GG_HIP(fill_async<int32_t>(ctr, 0, 2 * kXcdSlices, s));
    hipLaunchKernelGGL(k_slice_long<0>, dim3(grid_for(cells, 256, 16384)), dim3(256), 0, s,
                       g->sl_off[dir].as<int32_t>(), n, ctr, (const int32_t *)nullptr, (int32_t *)nullptr,
                       (int32_t *)nullptr);
    GG_HIP(hipGetLastError());
    int32_t h[2 * kXcdSlices];
    GG_HIP(hipMemcpyAsync(h, ctr, sizeof(h), hipMemcpyDeviceToHost, s));
    GG_HIP(hipStreamSynchronize(s));
    int32_t base[2 * (kXcdSlices + 1)];
    for (int c = 0; c < 2; c++) {
      base[c * (kXcdSlices + 1)] = 0;
      for (int t = 0; t < kXcdSlices; t++)
        base[c * (kXcdSlices + 1) + t + 1] = base[c * (kXcdSlices + 1) + t] + h[c * kXcdSlices + t];
    }
    const int32_t nmid = base[kXcdSlices], nlarge = base[2 * kXcdSlices + 1];
    GG_TRY(g->sl_long[dir].ensure(4 * (size_t)(nmid + nlarge + 1)));
    std::memcpy(g->sl_mid_base[dir], base, sizeof(int32_t) * (kXcdSlices + 1));
    std::memcpy(g->sl_large_base[dir], base + kXcdSlices + 1, sizeof(int32_t) * (kXcdSlices + 1));
    if (nmid + nlarge) {
      int32_t *dbase = ctr + 2 * kXcdSlices;
      GG_HIP(hipMemcpyAsync(dbase, base, sizeof(base), hipMemcpyHostToDevice, s));
      GG_HIP(fill_async<int32_t>(ctr, 0, 2 * kXcdSlices, s));
      hipLaunchKernelGGL(k_slice_long<1>, dim3(grid_for(cells, 256, 16384)), dim3(256), 0, s,
                         g->sl_off[dir].as<int32_t>(), n, ctr, (const int32_t *)dbase, g->sl_long[dir].as<int32_t>(),
                         g->sl_long[dir].as<int32_t>() + nmid);
      GG_HIP(hipGetLastError());
    }
    GG_HIP(hipStreamSynchronize(s));  // cnt / tmp are freed on return
    g->sl_ready[dir] = true;
    g->sl_val_version[dir] = -1;
  }
  // Slice-ordered transition weights (F32 graphs).
  if (g->value_dtype == GG_VALUES_F32 && g->sl_val_version[dir] != g->pr_weights_version) {
    GG_TRY(g->sl_val[dir].ensure(4 * (size_t)std::max<int64_t>(m, 1)));
    const float *p = g->pr[dir == 0 ? 3 : 2].as<float>();
    hipLaunchKernelGGL(k_gather_perm<float>, dim3(grid_for(m, 256, 16384)), dim3(256), 0, s, p,
                       g->sl_perm[dir].as<int32_t>(), m, g->sl_val[dir].as<float>());
    GG_HIP(hipGetLastError());
    g->sl_val_version[dir] = g->pr_weights_version;
  }
  GG_TRY(g->sl_part.ensure(8 * (size_t)cells));
  L->off = g->sl_off[dir].as<int32_t>();
  L->mid = g->sl_long[dir].as<int32_t>();
  L->large = L->mid + g->sl_mid_base[dir][kXcdSlices];
  L->n = n;
  std::memcpy(L->mid_base, g->sl_mid_base[dir], sizeof(L->mid_base));
  std::memcpy(L->large_base, g->sl_large_base[dir], sizeof(L->large_base));
  return GG_OK;
}

namespace {

// Device input: copy a host array into buf unless the flags say device.
int stage_input(gg_graph *g, const float *p, uint32_t flags, gg::DevBuf &buf, const float **out) {
  if (!p) {
    *out = nullptr;
    return GG_OK;
  }
  if (flags & GG_OUT_DEVICE) {
    *out = p;
    return GG_OK;
  }
  GG_TRY(buf.ensure(4 * (size_t)g->n));
  GG_HIP(hipMemcpyAsync(buf.p, p, 4 * (size_t)g->n, hipMemcpyHostToDevice, g->stream));
  *out = buf.as<float>();
  return GG_OK;
}

// Run sweeps (forward or adjoint op) to convergence; returns iterations.  The
// sweep is the merge-path segmented row sum over CSC (dir 0) or CSR (dir 1)
// rows or, with slices, the XCD-sliced partials + combine.
template <class Op>
int iterate(gg_graph *g, Op op, int dir, const SliceLists *slices, PrState *st, double tol, int32_t max_iter,
            const char *kname, double bytes_per_sweep, int32_t *iters_out, double *resid_out,
            double *scale_out = nullptr) {
  hipStream_t s = g->stream;
  const int32_t *done = &st->done;
  float2 *part = slices ? g->sl_part.as<float2>() : nullptr;
  const int32_t *offs = dir == 0 ? g->dg.column_offsets : g->dg.row_offsets;
  const int32_t *parts = nullptr;
  SegCarry *carry = nullptr;
  if (!slices) {
    GG_TRY(dir == 0 ? pull_partition(g, &parts) : csr_partition(g, &parts));
    GG_TRY(g->seg_carry.ensure(sizeof(SegCarry) * (size_t)seg_tiles(g->n, g->m)));
    carry = g->seg_carry.as<SegCarry>();
  }
  // Read back from PrState::scale on: scale, resid_last, iters, done.
  struct H {
    double scale, resid_last;
    int32_t iters, done;
  };
  static_assert(offsetof(PrState, resid_last) == offsetof(PrState, scale) + offsetof(H, resid_last) &&
                    offsetof(PrState, iters) == offsetof(PrState, scale) + offsetof(H, iters) &&
                    offsetof(PrState, done) == offsetof(PrState, scale) + offsetof(H, done),
                "PrState read-back layout");
  for (;;) {
    for (int j = 0; j < kItersPerCheck; j++) {
      {
        TimedRegion t(g->prof, kname);
        if (slices)
          slice_rows_launch(op, *slices, part, done, s);
        else
          seg_rows_launch(op, offs, g->n, g->m, parts, carry, done, g->opt.expand_grid, s);
      }
      hipLaunchKernelGGL(k_pr_finalize, dim3(1), dim3(1), 0, s, st, tol, max_iter);
    }
    GG_HIP(hipGetLastError());
    GG_HIP(hipMemcpyAsync(g->h_pinned, &st->scale, sizeof(H), hipMemcpyDeviceToHost, s));
    GG_HIP(hipStreamSynchronize(s));
    H h;
    std::memcpy(&h, g->h_pinned, sizeof(H));
    if (h.done) {
      *iters_out = h.iters;
      *resid_out = h.resid_last;
      if (scale_out) *scale_out = h.scale;
      g->prof.add_bytes(kname, bytes_per_sweep * h.iters);
      return GG_OK;
    }
  }
}

}  // namespace
}  // namespace gg

extern "C" {

int gg_pagerank(gg_graph *g, float damping, const float *teleport, double tol_l1, int32_t max_iter, float *rank,
                uint32_t flags, gg_stats *stats) {
  using namespace gg;
  gg::clear_error();
  GG_CHECK_ARG(g, "graph is NULL");
  GG_CHECK_ARG(damping >= 0.f && damping < 1.f, "damping must be in [0, 1)");
  GG_CHECK_ARG(tol_l1 >= 0 && max_iter >= 1, "need tol_l1 >= 0 and max_iter >= 1");
  GG_TRY(activate(g));
  const int32_t n = g->n;
  const int64_t m = g->m;
  hipStream_t s = g->stream;
  const bool weighted = g->value_dtype == GG_VALUES_F32;
  g->prof.enabled = (flags & GG_INSTRUMENT) != 0;
  GG_TRY(ensure_transition(g));
  GG_TRY(g->pr[4].ensure(8 * (size_t)n));  // r[2]
  GG_TRY(g->pr[5].ensure(8 * (size_t)n));  // x[2]
  GG_TRY(g->pr[6].ensure(sizeof(PrState)));
  const float *tele = nullptr;
  if (g->reordered && teleport)
    GG_TRY(reorder_in_vertices(g, teleport, flags, g->pr[7], reinterpret_cast<const void **>(&tele)));
  else
    GG_TRY(stage_input(g, teleport, flags, g->pr[7], &tele));
  float *r0 = g->pr[4].as<float>(), *r1 = r0 + n;
  float *x0 = g->pr[5].as<float>(), *x1 = x0 + n;
  PrState *st = g->pr[6].as<PrState>();

  ScopedEvents ev;
  GG_HIP(ev.create());
  hipEvent_t e0 = ev.e0, e1 = ev.e1;
  GG_HIP(hipEventRecord(e0, s));
  GG_HIP(fill_async<unsigned char>(reinterpret_cast<unsigned char *>(st), 0, sizeof(PrState), s));
  hipLaunchKernelGGL(k_pr_init, dim3(grid_for(n, kRowBlock, 4096)), dim3(kRowBlock), 0, s, r0,
                     weighted ? nullptr : x0, tele, g->pr[0].as<float>(), n, st, 0, (const float *)nullptr);
  hipLaunchKernelGGL(k_pr_start, dim3(1), dim3(1), 0, s, st, 0);
  GG_HIP(hipGetLastError());
  PrForwardOp op{};
  op.off = g->dg.column_offsets;
  op.src = g->dg.row_indices;
  op.pcsc = weighted ? g->pr[3].as<float>() : nullptr;
  op.inv_out = g->pr[0].as<float>();
  op.r[0] = r0;
  op.r[1] = r1;
  op.x[0] = x0;
  op.x[1] = x1;
  op.tele = tele;
  op.d = damping;
  op.uniform_tele = (float)((1.0 - damping) / (double)n);
  op.inv_n = 1.0 / (double)n;
  op.st = st;
  const double per_sweep = (weighted ? 16.0 * n + 12.0 * m : 20.0 * n + 8.0 * m) + (tele ? 4.0 * n : 0.0);
  SliceLists L{};
  const bool sliced = use_slices(g);
  if (sliced) {
    GG_TRY(ensure_slices(g, 0, &L));
    op.src = g->sl_idx[0].as<int32_t>();
    op.pcsc = weighted ? g->sl_val[0].as<float>() : nullptr;
  }
  int32_t iters = 0;
  double resid = 0;
  GG_TRY(iterate(g, op, 0, sliced ? &L : nullptr, st, tol_l1, max_iter, "pagerank.sweep", per_sweep, &iters, &resid));
  GG_HIP(hipEventRecord(e1, s));
  GG_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  finish_profile(g);
  const float *res = (iters & 1) ? r1 : r0;
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->iterations = iters;
    stats->reached_vertices = n;
    stats->reached_edges = m;
    stats->edges_inspected = (int64_t)iters * m;
    stats->alg_bytes = per_sweep * iters;
    stats->gpu_ms = ms;
    stats->residual = resid;
  }
  if (g->reordered)
    GG_TRY(reorder_out_vertices(g, res, rank, flags));
  else
    GG_TRY(copy_out(g, rank, res, 4 * (size_t)n, flags));
  if (resid > tol_l1) return set_error(GG_ERR_NOT_CONVERGED, "PageRank: residual %.3e > tol %.3e after %d sweeps",
                                       resid, tol_l1, iters);
  return GG_OK;
}

int gg_pagerank_adjoint(gg_graph *g, float damping, const float *rank, const float *c, double tol_l1,
                        int32_t max_iter, float *grad_w, float *grad_t, uint32_t flags, gg_stats *stats) {
  using namespace gg;
  gg::clear_error();
  GG_CHECK_ARG(g && rank && c, "graph/rank/c is NULL");
  GG_CHECK_ARG(damping >= 0.f && damping < 1.f, "damping must be in [0, 1)");
  GG_CHECK_ARG(tol_l1 >= 0 && max_iter >= 1, "need tol_l1 >= 0 and max_iter >= 1");
  // Int32 edge values (SSSP weights) do not enter PageRank: transitions are then 1/outdeg.
  GG_TRY(activate(g));
  const int32_t n = g->n;
  const int64_t m = g->m;
  hipStream_t s = g->stream;
  const bool weighted = g->value_dtype == GG_VALUES_F32;
  g->prof.enabled = (flags & GG_INSTRUMENT) != 0;
  GG_TRY(ensure_bins(g, 1));
  GG_TRY(ensure_transition(g));
  GG_TRY(g->batch[0].ensure(8 * (size_t)n));  // lambda[2]
  GG_TRY(g->batch[1].ensure(4 * (size_t)n));  // rowsum
  GG_TRY(g->batch[2].ensure(sizeof(PrState)));
  const float *d_rank = nullptr, *d_c = nullptr;
  if (g->reordered) {
    GG_TRY(reorder_in_vertices(g, rank, flags, g->batch[3], reinterpret_cast<const void **>(&d_rank)));
    GG_TRY(reorder_in_vertices(g, c, flags, g->batch[4], reinterpret_cast<const void **>(&d_c)));
  } else {
    GG_TRY(stage_input(g, rank, flags, g->batch[3], &d_rank));
    GG_TRY(stage_input(g, c, flags, g->batch[4], &d_c));
  }
  // Gradients go straight to the caller's device buffers unless they need
  // mapping back to the caller's order.
  const bool dev_out = (flags & GG_OUT_DEVICE) && !g->reordered;
  float *l0 = g->batch[0].as<float>(), *l1 = l0 + n;
  PrState *st = g->batch[2].as<PrState>();

  ScopedEvents ev;
  GG_HIP(ev.create());
  hipEvent_t e0 = ev.e0, e1 = ev.e1;
  GG_HIP(hipEventRecord(e0, s));
  GG_HIP(fill_async<unsigned char>(reinterpret_cast<unsigned char *>(st), 0, sizeof(PrState), s));
  hipLaunchKernelGGL(k_pr_init, dim3(grid_for(n, kRowBlock, 4096)), dim3(kRowBlock), 0, s, l0, (float *)nullptr,
                     (const float *)nullptr, g->pr[0].as<float>(), n, st, 1, d_c);
  hipLaunchKernelGGL(k_pr_start, dim3(1), dim3(1), 0, s, st, 1);
  GG_HIP(hipGetLastError());
  PrAdjointOp op{};
  op.off = g->dg.row_offsets;
  op.dst = g->dg.column_indices;
  op.pcsr = weighted ? g->pr[2].as<float>() : nullptr;
  op.inv_out = g->pr[0].as<float>();
  op.lam[0] = l0;
  op.lam[1] = l1;
  op.c = d_c;
  op.rowsum = g->batch[1].as<float>();
  op.d = damping;
  op.inv_n = 1.0 / (double)n;
  op.st = st;
  const double per_sweep = weighted ? 24.0 * n + 12.0 * m : 24.0 * n + 8.0 * m;
  SliceLists L{};
  const bool sliced = use_slices(g);
  if (sliced) {
    GG_TRY(ensure_slices(g, 1, &L));
    op.dst = g->sl_idx[1].as<int32_t>();
    op.pcsr = weighted ? g->sl_val[1].as<float>() : nullptr;
  }
  int32_t iters = 0;
  double resid = 0, scale = 1.0;  // scale = ||c||_1 (1 when c = 0)
  GG_TRY(iterate(g, op, 1, sliced ? &L : nullptr, st, tol_l1, max_iter, "pagerank.adjoint_sweep", per_sweep, &iters,
                 &resid, &scale));
  const float *lam = (iters & 1) ? l1 : l0;
  // Gradients.
  float *gw = nullptr, *gt = nullptr;
  if (grad_w) {
    if (dev_out) {
      gw = grad_w;
    } else {
      GG_TRY(g->batch[5].ensure(4 * (size_t)std::max<int64_t>(m, 1)));
      gw = g->batch[5].as<float>();
    }
    TimedRegion t(g->prof, "pagerank.gradient");
    const double *S = weighted ? g->pr[1].as<double>() : nullptr;
    const float *pcsr = weighted ? g->pr[2].as<float>() : nullptr;
    for (int b = 0; b < 3; b++) {
      const int32_t cnt = g->bin_count[1][b];
      if (!cnt) continue;
      if (b == 0)
        hipLaunchKernelGGL(k_pr_grad_rows, dim3(grid_for(cnt, kRowBlock, 4096)), dim3(kRowBlock), 0, s,
                           g->dg.row_offsets, g->dg.column_indices, pcsr, S, d_rank, lam, damping,
                           g->bins[1][b].as<int32_t>(), cnt, gw);
      else
        hipLaunchKernelGGL(k_pr_grad_rows_wave, dim3(grid_for((int64_t)cnt * kWave, kRowBlock, 4096)), dim3(kRowBlock),
                           0, s, g->dg.row_offsets, g->dg.column_indices, pcsr, S, d_rank, lam, damping,
                           g->bins[1][b].as<int32_t>(), cnt, gw);
    }
    GG_HIP(hipGetLastError());
    g->prof.add_bytes("pagerank.gradient", 16.0 * m + 16.0 * n);
  }
  if (grad_t) {
    if (dev_out) {
      gt = grad_t;
    } else {
      GG_TRY(g->batch[6].ensure(4 * (size_t)n));
      gt = g->batch[6].as<float>();
    }
    hipLaunchKernelGGL(k_scale, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, lam, 1.0f - damping, gt, n);
    GG_HIP(hipGetLastError());
  }
  GG_HIP(hipEventRecord(e1, s));
  GG_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  finish_profile(g);
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->iterations = iters;
    stats->reached_vertices = n;
    stats->reached_edges = m;
    stats->edges_inspected = (int64_t)(iters + 1) * m;
    stats->alg_bytes = per_sweep * iters + (grad_w ? 16.0 * m + 16.0 * n : 0.0);
    stats->gpu_ms = ms;
    stats->residual = resid;
  }
  if (g->reordered) {
    if (grad_w) GG_TRY(reorder_out_edges(g, gw, grad_w, flags));
    if (grad_t) GG_TRY(reorder_out_vertices(g, gt, grad_t, flags));
  } else {
    if (grad_w && !dev_out) GG_TRY(copy_out(g, grad_w, gw, 4 * (size_t)m, flags));
    if (grad_t && !dev_out) GG_TRY(copy_out(g, grad_t, gt, 4 * (size_t)n, flags));
  }
  if (resid > tol_l1 * scale)
    return set_error(GG_ERR_NOT_CONVERGED, "adjoint: residual %.3e > tol %.3e x ||c||_1 after %d sweeps", resid, tol_l1,
                     iters);
  return GG_OK;
}

}  // extern "C"
