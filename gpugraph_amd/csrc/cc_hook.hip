// Connected components by hooking + pointer jumping (GG_CC_HOOK).
//
// Gives the same labels as the reference's label propagation (Algorithms/CC/
// cc.h, checked against CPUCC cc.cu:127-169): the minimum vertex id of each
// component.  Every hook links the larger of two roots under the smaller one
// with an atomicCAS, so parent[x] <= x always holds and the root of each
// final tree is the component's minimum id.  Two passes over the edges
// instead of one pass per unit of diameter, so it suits high-diameter graphs
// (lattices) where label propagation needs thousands of iterations.
// Symmetric (undirected) graphs only.
//
// Byte model: hook 8n + 8m (offsets, parents, column ids, parent probes of
// the m/2 kept edges), compress 8n.
#include <hip/hip_runtime.h>

#include <cstring>

#include "internal.h"

namespace gg {
namespace {

using namespace GASengine;

__device__ __forceinline__ int32_t find_root(int32_t *parent, int32_t x) {
  for (;;) {
    const int32_t p = __atomic_load_n(&parent[x], __ATOMIC_RELAXED);
    if (p == x) return x;
    const int32_t gp = __atomic_load_n(&parent[p], __ATOMIC_RELAXED);
    if (gp != p) __atomic_store_n(&parent[x], gp, __ATOMIC_RELAXED);  // path halving: gp <= p keeps the order
    x = gp;
  }
}

__device__ __forceinline__ void hook(int32_t *parent, int32_t u, int32_t v) {
  int32_t ru = find_root(parent, u), rv = find_root(parent, v);
  while (ru != rv) {
    if (ru > rv) {
      const int32_t t = ru;
      ru = rv;
      rv = t;
    }
    const int32_t old = atomicCAS(&parent[rv], rv, ru);
    if (old == rv) break;
    rv = find_root(parent, old);
    ru = find_root(parent, ru);
  }
}

// Low-degree vertices: one thread each; degree >= 64: one wave.
__global__ void __launch_bounds__(256)
k_cc_hook(const int32_t *__restrict__ ro, const int32_t *__restrict__ col, int32_t *parent, int32_t n) {
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = tid - lane_id(); base < n; base += stride) {
    const int32_t u = (int32_t)(base + lane_id());
    int32_t b = 0, e = 0;
    if (u < n) {
      b = ro[u];
      e = ro[u + 1];
    }
    const bool big = e - b >= kWave;
    if (!big)
      for (int32_t k = b; k < e; k++) {
        const int32_t v = col[k];
        if (v < u) hook(parent, u, v);
      }
    uint64_t mask = ballot(big);
    while (mask) {
      const int l = __builtin_ctzll(mask);
      mask &= mask - 1;
      const int32_t wu = __shfl(u, l), wb = __shfl(b, l), we = __shfl(e, l);
      for (int32_t k = wb + lane_id(); k < we; k += kWave) {
        const int32_t v = col[k];
        if (v < wu) hook(parent, wu, v);
      }
    }
  }
}

__global__ void k_cc_compress(int32_t *parent, int32_t n) {
  for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    parent[v] = find_root(parent, v);
}

}  // namespace

int cc_hooking(gg_graph *g, int32_t *label_out, uint32_t flags, gg_stats *stats) {
  GG_TRY(activate(g));
  if (g->directed) return set_error(GG_ERR_UNSUPPORTED, "GG_CC_HOOK needs a symmetric (undirected) graph");
  const int32_t n = g->n;
  hipStream_t s = g->stream;
  int32_t *parent;
  if (flags & GG_OUT_DEVICE) {
    GG_CHECK_ARG(label_out, "label is NULL");
    parent = label_out;
  } else {
    GG_TRY(g->vtx[1].ensure(4 * (size_t)n));
    parent = g->vtx[1].as<int32_t>();
  }
  g->prof.enabled = (flags & GG_INSTRUMENT) != 0;
  hipEvent_t e0, e1;
  GG_HIP(hipEventCreate(&e0));
  GG_HIP(hipEventCreate(&e1));
  GG_HIP(hipEventRecord(e0, s));
  {
    TimedRegion t(g->prof, "cc.hook");
    GG_HIP(sequence_async<int32_t>(parent, n, s));
    hipLaunchKernelGGL(k_cc_hook, dim3(grid_for(n, 256, 16384)), dim3(256), 0, s, g->dg.row_offsets,
                       g->dg.column_indices, parent, n);
    GG_HIP(hipGetLastError());
  }
  {
    TimedRegion t(g->prof, "cc.compress");
    hipLaunchKernelGGL(k_cc_compress, dim3(grid_for(n, 256, 16384)), dim3(256), 0, s, parent, n);
    GG_HIP(hipGetLastError());
  }
  GG_HIP(hipEventRecord(e1, s));
  GG_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  g->prof.add_bytes("cc.hook", 8.0 * n + 8.0 * (double)g->m);
  g->prof.add_bytes("cc.compress", 8.0 * n);
  finish_profile(g);
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->iterations = 2;
    stats->edges_inspected = g->m;
    stats->alg_bytes = 16.0 * n + 8.0 * (double)g->m;
    stats->gpu_ms = ms;
    GG_TRY(reach_stats(g, parent, INT32_MIN, stats));
  }
  if (flags & GG_OUT_DEVICE) return GG_OK;
  return copy_out(g, label_out, parent, 4 * (size_t)n, flags);
}

}  // namespace gg
