// C ABI runs of the four vertex programs through the generic GAS engine
// (EnactorVertexCentric<CsrProblem<Program>, Program>):
//   gg_sssp (default path), gg_cc (label propagation), gg_bfs with
//   GG_BFS_PUSH_ONLY, gg_pagerank_reference.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstring>
#include <string>

#include "internal.h"
#include "programs/bfs.h"
#include "programs/cc.h"
#include "programs/pagerank.h"
#include "programs/sssp.h"
#include "programs/variants.h"

namespace gg {

// Times the enactor's kernel classes under "<program>.<kernel>".
struct PrefixTimer : GASengine::KernelTimer {
  EventProfiler &p;
  std::string prefix;
  std::string cur;
  PrefixTimer(EventProfiler &p_, const char *pre) : p(p_), prefix(pre) {}
  void begin(const char *name) override {
    cur = prefix + "." + name;
    p.begin(cur.c_str());
  }
  void end(const char *name) override {
    cur = prefix + "." + name;
    p.end(cur.c_str());
  }
};

template <class P>
struct Slot;
template <>
struct Slot<bfs> {
  static constexpr int index = 0;
  static constexpr const char *name = "bfs_gas";
  static constexpr int default_threshold = 0;  // Algorithms/BFS/register.h:41
};
template <>
struct Slot<sssp> {
  static constexpr int index = 1;
  static constexpr const char *name = "sssp";
  static constexpr int default_threshold = 10000;  // Algorithms/SSSP/register.h:39
};
template <>
struct Slot<cc> {
  static constexpr int index = 2;
  static constexpr const char *name = "cc";
  static constexpr int default_threshold = 10000;  // Algorithms/CC/register.h:39
};
template <>
struct Slot<pagerank> {
  static constexpr int index = 3;
  static constexpr const char *name = "pagerank_ref";
  static constexpr int default_threshold = 1000;  // Algorithms/PageRank/register.h:40
};

template <>
struct Slot<bfs_reverse> {
  static constexpr int index = 4;
  static constexpr const char *name = "bfs_reverse";
  static constexpr int default_threshold = 0;
};
template <>
struct Slot<cc_weak> {
  static constexpr int index = 5;
  static constexpr const char *name = "cc_weak";
  static constexpr int default_threshold = 10000;
};
template <>
struct Slot<pagerank_all> {
  static constexpr int index = 6;
  static constexpr const char *name = "pagerank_all";
  static constexpr int default_threshold = 1000;
};

template <class P>
using Problem = GASengine::CsrProblem<P>;

template <class P>
int get_problem(gg_graph *g, Problem<P> **out) {
  auto &slot = g->prog_state[Slot<P>::index];
  if (!slot) {
    auto *p = new (std::nothrow) Problem<P>();
    if (!p) return set_error(GG_ERR_OOM, "host allocation failed");
    hipError_t e = p->Attach(g->dg, g->stream);
    if (e != hipSuccess) {
      delete p;
      return set_error(e == hipErrorOutOfMemory ? GG_ERR_OOM : GG_ERR_HIP, "workspace allocation: %s",
                       hipGetErrorString(e));
    }
    slot = std::shared_ptr<void>(p, [](void *q) {
      auto *pp = static_cast<Problem<P> *>(q);
      P::Release(pp->vertex_list, pp->edge_list);
      delete pp;
    });
  }
  auto *p = static_cast<Problem<P> *>(slot.get());
  p->stream = g->stream;
  *out = p;
  return GG_OK;
}

// Algorithmic bytes of iterations with frontier F, edges E, next frontier N
// (per iteration, or summed over a run: the model is linear).
template <class P>
double iteration_bytes(double F, double E, double N) {
  if (std::is_same<P, sssp>::value) return 16 * F + 12 * E + 8 * N;
  if (std::is_same<P, cc>::value || std::is_same<P, cc_weak>::value) return 16 * F + 8 * E + 8 * N;
  if (std::is_same<P, bfs>::value || std::is_same<P, bfs_reverse>::value) return 12 * F + 8 * E + 8 * N;
  return 12 * F + 12 * E + 8 * N;  // pagerank: gather reads ri + rank + outdeg per in-edge
}

template <class P>
double iteration_bytes(const GASengine::IterationRecord &r) {
  return iteration_bytes<P>(r.frontier, (double)r.edges, r.next);
}

template <class P>
int run_program(gg_graph *g, int num_srcs, int *srcs, typename P::DataType *out, uint32_t flags, gg_stats *stats,
                typename P::DataType sentinel, bool count_reach) {
  GG_TRY(activate(g));
  Problem<P> *prob = nullptr;
  GG_TRY(get_problem<P>(g, &prob));
  GASengine::EnactorVertexCentric<Problem<P>, P> enactor;
  enactor.config.wave_degree = g->opt.wave_degree;
  enactor.config.block_degree = g->opt.block_degree;
  enactor.config.expand_grid = g->opt.gas_expand_grid;
  enactor.config.expand_vt = g->opt.lbs_tile / GASengine::kExpandBlock;
  // Under the per-algorithm default threshold, a small frontier holding a
  // row longer than 4 * block_degree (a hub that one tiered workgroup would
  // walk alone) is expanded by LBS; an explicit threshold decides alone.
  enactor.config.tier_max_degree = g->opt.threshold < 0 ? 4 * g->opt.block_degree : 0;
  g->prof.enabled = (flags & GG_INSTRUMENT) != 0;
  PrefixTimer timer(g->prof, Slot<P>::name);
  if (g->prof.enabled) enactor.timer = &timer;
  const int threshold = g->opt.threshold >= 0 ? g->opt.threshold : Slot<P>::default_threshold;

  ScopedEvents ev;
  GG_HIP(ev.create());
  hipEvent_t e0 = ev.e0, e1 = ev.e1;
  GG_HIP(hipEventRecord(e0, g->stream));
  hipError_t err = enactor.EnactIterativeSearch(*prob, nullptr, g->directed, num_srcs, srcs, g->opt.iter_num,
                                                threshold);
  (void)hipEventRecord(e1, g->stream);
  (void)hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  if (err != hipSuccess)
    return set_error(err == hipErrorOutOfMemory ? GG_ERR_OOM : GG_ERR_HIP, "%s: %s", Slot<P>::name,
                     hipGetErrorString(err));
  typename P::DataType *res = P::resultPointer(prob->vertex_list);
  if (g->prof.enabled) {
    for (auto &r : enactor.history) {
      const double b = iteration_bytes<P>(r);
      const std::string base = std::string(Slot<P>::name) + ".";
      // Edge traffic to the expand kernel, per-vertex traffic to the count pass.
      const double fb = b - (std::is_same<P, cc>::value ? 8 : 12) * (double)r.edges - 8.0 * r.next;
      g->prof.add_bytes((base + (r.lbs ? "lbs_expand" : "tiered_expand")).c_str(), b - fb);
      g->prof.add_bytes((base + (r.lbs ? "lbs_count" : "prologue")).c_str(), fb);
    }
    finish_profile(g);
  }
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->iterations = enactor.iterations_run;
    stats->edges_inspected = enactor.total_edges;
    stats->gpu_ms = ms;
    stats->alg_bytes = iteration_bytes<P>((double)enactor.frontier_total, (double)enactor.total_edges,
                                          (double)enactor.next_total);
    if (count_reach) GG_TRY(reach_stats(g, (const int32_t *)res, *(int32_t *)&sentinel, stats));
  }
  return copy_out(g, out, res, sizeof(typename P::DataType) * (size_t)g->n, flags);
}

}  // namespace gg

namespace gg {
namespace {

// BFS tree parents from the depths: one wave per vertex scans its in-edges
// (CSC, sources ascending) 64 at a time; the first source one level up is the
// parent (the smallest such id).  pred[src] = src, -1 where unreached.  With
// old_of (a renumbered graph) the parent is the source with the smallest
// original id, and pred holds original ids.
__global__ void __launch_bounds__(256)
k_bfs_parents(const int32_t *__restrict__ co, const int32_t *__restrict__ ri, const int32_t *__restrict__ depth,
              int32_t n, int32_t src, const int32_t *__restrict__ old_of, int32_t *__restrict__ pred) {
  const int64_t waves = (int64_t)gridDim.x * (256 / GASengine::kWave);
  const int lane = GASengine::lane_id();
  for (int64_t v = (int64_t)blockIdx.x * (256 / GASengine::kWave) + threadIdx.x / GASengine::kWave; v < n;
       v += waves) {
    const int32_t dv = depth[v];
    int32_t p = -1;
    if (v == src) {
      p = old_of ? old_of[src] : src;
    } else if (dv > 0) {
      const int32_t b = co[v], e = co[v + 1];
      int32_t best = INT_MAX;
      for (int32_t k0 = b; k0 < e; k0 += GASengine::kWave) {
        const int32_t k = k0 + lane;
        const int32_t u = k < e ? ri[k] : -1;
        const bool up = k < e && depth[u] == dv - 1;
        if (!old_of) {
          const uint64_t hit = GASengine::ballot(up);
          if (hit) {
            best = __shfl(u, __builtin_ctzll(hit));
            break;
          }
        } else if (up) {
          best = min(best, old_of[u]);
        }
      }
      if (old_of)
        for (int off = 32; off > 0; off >>= 1) best = min(best, __shfl_xor(best, off));
      p = best == INT_MAX ? -1 : best;
    }
    if (lane == 0) pred[v] = p;
  }
}

}  // namespace

// Cores of the public entry points: internal vertex ids, any output pointer.
int bfs_direction_optimizing(gg_graph *g, int32_t src, int32_t *depth, uint32_t flags, gg_stats *stats);
int cc_hooking(gg_graph *g, int32_t *label, uint32_t flags, gg_stats *stats);
int cc_label_propagation(gg_graph *g, int32_t *label, uint32_t flags, gg_stats *stats);
int sssp_near_far(gg_graph *g, int32_t src, int32_t *dist, uint32_t flags, gg_stats *stats);

int bfs_core(gg_graph *g, int32_t src, int32_t *depth, uint32_t flags, gg_stats *stats) {
  int s = src;
  if (flags & GG_BFS_REVERSE) return run_program<bfs_reverse>(g, 1, &s, depth, flags, stats, -1, true);
  if (!(flags & GG_BFS_PUSH_ONLY)) return bfs_direction_optimizing(g, src, depth, flags, stats);
  return run_program<bfs>(g, 1, &s, depth, flags, stats, -1, true);
}

int sssp_core(gg_graph *g, int32_t src, int32_t *dist, uint32_t flags, gg_stats *stats) {
  int s = src;
  if (!(flags & GG_SSSP_GAS)) return sssp_near_far(g, src, dist, flags, stats);
  return run_program<sssp>(g, 1, &s, dist, flags, stats, sssp::INIT_VALUE, true);
}

int cc_core(gg_graph *g, int32_t *label, uint32_t flags, gg_stats *stats) {
  if (flags & GG_CC_WEAK) {
    if (!g->directed) return set_error(GG_ERR_INVALID, "GG_CC_WEAK is for directed graphs (use gg_cc without it)");
    return run_program<cc_weak>(g, 0, nullptr, label, flags, stats, INT_MIN, true);
  }
  if (flags & GG_CC_HOOK) return cc_hooking(g, label, flags, stats);
  if (!(flags & GG_CC_GAS)) return cc_label_propagation(g, label, flags, stats);
  // Every vertex is labelled: reach_stats counts all of them.
  return run_program<cc>(g, 0, nullptr, label, flags, stats, INT_MIN, true);
}

// A renumbered graph runs a core into re_out[0] and maps the result back.
// The mapping is armed as the core's hand-off: a device-driven core enqueues
// it from its own result buffer ahead of its last read-back (one
// synchronisation for both); otherwise it runs here.
template <class Core>
int with_mapped_output(gg_graph *g, void *out, uint32_t flags, Core core) {
  GG_TRY(activate(g));
  GG_TRY(g->re_out[0].ensure(4 * (size_t)g->n));
  g->handoff = {out, flags, out != nullptr, false};
  const int rc = core(g->re_out[0].as<int32_t>(), flags | GG_OUT_DEVICE);
  const bool done = g->handoff.done;
  g->handoff = {};
  GG_TRY(rc);
  if (done) return GG_OK;
  return reorder_out_vertices(g, g->re_out[0].p, out, flags);
}

}  // namespace gg

extern "C" {

int gg_bfs(gg_graph *g, int32_t src, int32_t *depth, uint32_t flags, gg_stats *stats) {
  using namespace gg;
  gg::clear_error();
  GG_CHECK_ARG(g, "graph is NULL");
  GG_CHECK_ARG(src >= 0 && src < g->n, "src %d out of range [0, %d)", src, g->n);
  if (!g->reordered) return bfs_core(g, src, depth, flags, stats);
  const int32_t s = map_vertex(g, src);
  return with_mapped_output(g, depth, flags,
                            [&](int32_t *o, uint32_t f) { return bfs_core(g, s, o, f, stats); });
}

int gg_bfs_tree(gg_graph *g, int32_t src, int32_t *depth, int32_t *pred, uint32_t flags, gg_stats *stats) {
  using namespace gg;
  gg::clear_error();
  GG_CHECK_ARG(g && pred, "graph/pred is NULL");
  GG_CHECK_ARG(src >= 0 && src < g->n, "src %d out of range [0, %d)", src, g->n);
  GG_CHECK_ARG(!(flags & GG_BFS_REVERSE), "gg_bfs_tree builds trees of forward BFS only");
  GG_TRY(activate(g));
  const bool dev = (flags & GG_OUT_DEVICE) && !g->reordered;
  const size_t bytes = 4 * (size_t)g->n;
  int32_t *d_depth = depth;
  if (!dev || !depth) {
    GG_TRY(g->tree[0].ensure(bytes));
    d_depth = g->tree[0].as<int32_t>();
  }
  const int32_t s = map_vertex(g, src);
  GG_TRY(bfs_core(g, s, d_depth, flags | GG_OUT_DEVICE, stats));
  int32_t *d_pred = pred;
  if (!dev) {
    GG_TRY(g->tree[1].ensure(bytes));
    d_pred = g->tree[1].as<int32_t>();
  }
  {
    TimedRegion t(g->prof, "bfs.parents");
    hipLaunchKernelGGL(k_bfs_parents, dim3(GASengine::grid_for((int64_t)g->n * GASengine::kWave, 256, 16384)),
                       dim3(256), 0, g->stream, g->dg.column_offsets, g->dg.row_indices, d_depth, g->n, s,
                       g->reordered ? g->re_order.as<int32_t>() : (const int32_t *)nullptr, d_pred);
    GG_HIP(hipGetLastError());
  }
  if (g->reordered) {
    GG_TRY(reorder_out_vertices(g, d_depth, depth, flags));
    GG_TRY(reorder_out_vertices(g, d_pred, pred, flags));
  } else {
    if (!dev && depth) GG_TRY(copy_out(g, depth, d_depth, bytes, flags));
    GG_TRY(copy_out(g, pred, d_pred, bytes, flags));
  }
  if (flags & GG_INSTRUMENT) finish_profile(g);
  return GG_OK;
}

int gg_sssp(gg_graph *g, int32_t src, int32_t *dist, uint32_t flags, gg_stats *stats) {
  using namespace gg;
  gg::clear_error();
  GG_CHECK_ARG(g, "graph is NULL");
  GG_CHECK_ARG(src >= 0 && src < g->n, "src %d out of range [0, %d)", src, g->n);
  GG_CHECK_ARG(g->value_dtype == GG_VALUES_I32, "SSSP needs int32 edge weights");
  GG_TRY(activate(g));
  GG_TRY(require_nonnegative_weights(g, "SSSP"));
  if (!g->reordered) return sssp_core(g, src, dist, flags, stats);
  const int32_t s = map_vertex(g, src);
  return with_mapped_output(g, dist, flags,
                            [&](int32_t *o, uint32_t f) { return sssp_core(g, s, o, f, stats); });
}

int gg_cc(gg_graph *g, int32_t *label, uint32_t flags, gg_stats *stats) {
  using namespace gg;
  gg::clear_error();
  GG_CHECK_ARG(g, "graph is NULL");
  if (!g->reordered) return cc_core(g, label, flags, stats);
  GG_TRY(activate(g));
  GG_TRY(g->re_out[0].ensure(4 * (size_t)g->n));
  GG_TRY(cc_core(g, g->re_out[0].as<int32_t>(), flags | GG_OUT_DEVICE, stats));
  return reorder_out_cc(g, g->re_out[0].as<int32_t>(), label, flags);
}

int gg_pagerank_reference(gg_graph *g, float *rank, uint32_t flags, gg_stats *stats) {
  using namespace gg;
  gg::clear_error();
  GG_CHECK_ARG(g, "graph is NULL");
  const bool all = flags & GG_PR_ALL_EDGES;
  GG_CHECK_ARG(!all || g->directed, "GG_PR_ALL_EDGES is for directed graphs");
  auto run = [&](float *o, uint32_t f) {
    return all ? run_program<pagerank_all>(g, 0, nullptr, o, f, stats, 0.0f, false)
               : run_program<pagerank>(g, 0, nullptr, o, f, stats, 0.0f, false);
  };
  if (!g->reordered) return run(rank, flags);
  return with_mapped_output(g, rank, flags,
                            [&](int32_t *o, uint32_t f) { return run(reinterpret_cast<float *>(o), f); });
}

}  // extern "C"
