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

// Hooking as an operator of GASengine::k_pull_reduce: the merge-path tiles of
// (vertices, adjacency) give every workgroup the same number of edges however
// skewed the degrees (hub rows spread over many workgroups), and each edge
// (v, u) with u < v hooks once.  No reduction: the LDS accumulator is unused.
struct HookOp {
  const int32_t *col;
  int32_t *parent;

  __device__ __forceinline__ int32_t identity() { return 0; }
  __device__ __forceinline__ int32_t edge(int32_t v, int32_t slot) {
    const int32_t u = col[slot];
    if (u < v) hook(parent, v, u);
    return 0;
  }
  __device__ __forceinline__ void combine(int32_t *, int32_t) {}
  __device__ __forceinline__ void item_done(int32_t, int32_t, bool) {}
  __device__ __forceinline__ void finish() {}
};

// Root without path halving: follows parents only.
__device__ __forceinline__ int32_t walk_to_root(const int32_t *parent, int32_t x) {
  for (;;) {
    const int32_t p = __atomic_load_n(&parent[x], __ATOMIC_RELAXED);
    if (p == x) return x;
    x = p;
  }
}

// Final labels: parent[v] = root of v.  Each parent[v] is written by v's
// thread only, with the root, and read by other threads' walks, which accept
// an old ancestor or the root alike.  (With path halving here, a walk that
// read parent[v] before v's thread stored the root could store an older
// ancestor after it, leaving a non-root label.)
__global__ void k_cc_compress(int32_t *parent, int32_t n) {
  for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    const int32_t r = walk_to_root(parent, v);
    __atomic_store_n(&parent[v], r, __ATOMIC_RELAXED);
  }
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
  const int32_t *parts = nullptr;
  GG_TRY(pull_partition(g, &parts));  // merge-path tiles over (vertices, adjacency), cached per graph
  g->prof.enabled = (flags & GG_INSTRUMENT) != 0;
  ScopedEvents ev;
  GG_HIP(ev.create());
  hipEvent_t e0 = ev.e0, e1 = ev.e1;
  GG_HIP(hipEventRecord(e0, s));
  {
    TimedRegion t(g->prof, "cc.hook");
    GG_HIP(sequence_async<int32_t>(parent, n, s));
    hipLaunchKernelGGL((k_pull_reduce<HookOp, int32_t>), dim3(g->opt.expand_grid), dim3(kExpandBlock), 0, s,
                       HookOp{g->dg.column_indices, parent}, g->dg.row_offsets, n, g->m, parts,
                       (const int32_t *)nullptr);
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
