// EnactorVertexCentric: the GAS iteration driver (reference:
// GASengine/enactor_vertex_centric.cuh:96-5072; its single-GPU sequence is
// gather -> apply -> post_apply -> expand -> contract, :3454-3771).
//
// Per iteration on the problem's stream:
//   [gather]  k_gather over in/out edges of the frontier    (gatherOverEdges)
//   if |F| > threshold  (LBS):
//     k_lbs_count  (apply + post_apply fused as the per-item prologue)
//     k_lbs_scan
//     k_lbs_expand (expand_edge + contract + dedup + enqueue per edge)
//   else (tiered):
//     k_frontier_prologue (apply + post_apply), if the program has them
//     k_tiered_expand
//   read back |F_next| (one 4-byte D2H per iteration, as the reference does
//   at enactor_vertex_centric.cuh:3801-3804), stop at 0 or iter_num.
//
// post_apply runs right after apply in the same thread: the barrier the
// reference places between them (kernel boundary) only matters for programs
// whose post_apply reads another vertex's apply output, and none of the
// vertex programs of this path do.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <vector>

#include "csr_problem.h"

namespace GASengine {

// Optional per-kernel timing hook (the reference's INSTRUMENT flag collects
// per-CTA clocks, b40c/util/kernel_runtime_stats.cuh:40-130).
struct KernelTimer {
  virtual void begin(const char *name) = 0;
  virtual void end(const char *name) = 0;
  virtual ~KernelTimer() = default;
};

struct EnactorConfig {
  int32_t wave_degree = 64;
  int32_t block_degree = 4096;
  int32_t expand_grid = 2048;  // persistent workgroups for k_lbs_expand
  int32_t expand_vt = kExpandVt;  // LBS tile = kExpandBlock * expand_vt positions (2, 4 or 8)
};

// Per-iteration record: frontier size in, edges expanded, frontier size out.
struct IterationRecord {
  int32_t frontier;
  int64_t edges;
  int32_t next;
  int lbs;
};

template <class CsrProblemT, class Program, bool INSTRUMENT = false>
class EnactorVertexCentric {
 public:
  EnactorConfig config;
  KernelTimer *timer = nullptr;
  std::vector<IterationRecord> history;
  int64_t total_edges = 0;
  int verbose = 0;

  explicit EnactorVertexCentric(bool verbose_ = false) : verbose(verbose_) {}

  // Reference 7-argument form (as called by Algorithms/SSSP/sssp.cu:397-398).
  hipError_t EnactIterativeSearch(CsrProblemT &problem, const int32_t * /*h_row_offsets*/, int directed,
                                  int num_srcs, int *srcs, int iter_num = INT_MAX, int threshold = 0) {
    hipStream_t s = problem.stream;
    Workspace &ws = problem.ws;
    DeviceGraph &g = problem.graph;
    const int32_t n = g.nodes;
    history.clear();
    total_edges = 0;

    GAS_RETURN_IF(problem.Reset());
    GAS_RETURN_IF(ws.AllocateValues());
    int *keys[3] = {ws.frontier[0], ws.frontier[1], ws.frontier[1]};
    typename Program::MiscType *vals[3];
    for (int i = 0; i < 3; i++) vals[i] = (typename Program::MiscType *)ws.frontier_values[i];
    engine_stream() = s;
    Program::Initialize(directed, n, (int)g.edges, num_srcs, srcs, (int *)g.row_offsets, (int *)g.column_indices,
                        (int *)g.column_offsets, (int *)g.row_indices, (typename Program::DataType *)g.edge_values,
                        problem.vertex_list, problem.edge_list, keys, vals);
    GAS_RETURN_IF(hipGetLastError());
    int32_t nf = Program::srcVertex() == SINGLE ? num_srcs : n;
    GAS_RETURN_IF(fill_async<int32_t>(ws.sizes, nf, 1, s));

    GasState<Program> st;
    st.g = g;
    st.vl = problem.vertex_list;
    st.el = problem.edge_list;
    st.gather_tmp = (typename Program::GatherType *)ws.gather_tmp;
    st.changed = ws.changed;
    st.do_apply = Program::applyOverEdges() != NO_APPLY_VERTICES;
    st.do_post = Program::postApplyOverEdges() != NO_POST_APPLY_VERTICES;
    const ExpandEdges ex = Program::expandOverEdges();
    const GatherEdges ga = Program::gatherOverEdges();

    for (int it = 0; it < iter_num && nf > 0; it++) {
      const int sel = it & 1;
      int32_t *f_in = ws.frontier[sel], *f_out = ws.frontier[sel ^ 1];
      int32_t *sz_in = ws.sizes + sel, *sz_out = ws.sizes + (sel ^ 1);
      st.iteration = it;
      st.dedup_next = ws.dedup[sel];
      st.dedup_prev = ws.dedup[sel ^ 1];
      GAS_RETURN_IF(hipMemsetAsync(sz_out, 0, sizeof(int32_t), s));

      if (ga != NO_GATHER_EDGES) {
        const int grid = grid_for((int64_t)nf * kWave, 256, 16384);
        if (ga == GATHER_IN_EDGES || ga == GATHER_ALL_EDGES) {
          tb("gather");
          hipLaunchKernelGGL((k_gather<Program, false>), dim3(grid), dim3(256), 0, s, st, f_in, sz_in, 1);
          te("gather");
        }
        if (ga == GATHER_OUT_EDGES || ga == GATHER_ALL_EDGES) {
          tb("gather");
          if (ga == GATHER_ALL_EDGES)
            hipLaunchKernelGGL((k_gather<Program, true>), dim3(grid), dim3(256), 0, s, st, f_in, sz_in, 0);
          else
            hipLaunchKernelGGL((k_gather<Program, false>), dim3(grid), dim3(256), 0, s, st, f_in, sz_in, 0);
          te("gather");
        }
        GAS_RETURN_IF(hipGetLastError());
      }

      int passes[2];
      int np = 0;
      if (ex == EXPAND_OUT_EDGES || ex == EXPAND_ALL_EDGES) passes[np++] = 0;
      if (ex == EXPAND_IN_EDGES || ex == EXPAND_ALL_EDGES) passes[np++] = 1;
      const bool lbs = nf > threshold;
      if (np == 0) {
        // No expansion: run the vertex phase only; the frontier empties.
        tb("prologue");
        hipLaunchKernelGGL(k_frontier_prologue<ProgramOp<Program>>, dim3(grid_for(nf, 256)), dim3(256), 0, s,
                           ProgramOp<Program>{st}, f_in, sz_in);
        te("prologue");
      }
      for (int p = 0; p < np; p++) {
        st.in_edges = passes[p];
        ProgramOp<Program> op{st};
        const bool first = p == 0;
        if (lbs) {
          static const char *names[4] = {"lbs_count", "lbs_scan", "lbs_partition", "lbs_expand"};
          LbsBuffers b{ws.scan_local, ws.partials, ws.info, ws.parts, ws.counters, config.expand_grid,
                       config.expand_vt};
          lbs_launch(op, first, f_in, sz_in, nf, (int64_t)n + g.edges, b, f_out, sz_out, s, [&](int phase) {
            if (phase > 0) te(names[phase - 1]);
            if (phase < 4) tb(names[phase]);
          });
        } else {
          if (first && (st.do_apply || st.do_post)) {
            tb("prologue");
            hipLaunchKernelGGL(k_frontier_prologue<ProgramOp<Program>>, dim3(grid_for(nf, 256)), dim3(256), 0, s,
                               op, f_in, sz_in);
            te("prologue");
            tb("tiered_expand");
            hipLaunchKernelGGL((k_tiered_expand<ProgramOp<Program>, false>), dim3(grid_for(nf, kExpandBlock)),
                               dim3(kExpandBlock), 0, s, op, f_in, sz_in, f_out, sz_out, config.wave_degree,
                               config.block_degree, ws.counters);
            te("tiered_expand");
          } else {
            // No apply phase: the prologue only clears dedup_prev, safe to fuse.
            tb("tiered_expand");
            if (first)
              hipLaunchKernelGGL((k_tiered_expand<ProgramOp<Program>, true>), dim3(grid_for(nf, kExpandBlock)),
                                 dim3(kExpandBlock), 0, s, op, f_in, sz_in, f_out, sz_out, config.wave_degree,
                                 config.block_degree, ws.counters);
            else
              hipLaunchKernelGGL((k_tiered_expand<ProgramOp<Program>, false>), dim3(grid_for(nf, kExpandBlock)),
                                 dim3(kExpandBlock), 0, s, op, f_in, sz_in, f_out, sz_out, config.wave_degree,
                                 config.block_degree, ws.counters);
            te("tiered_expand");
          }
        }
        GAS_RETURN_IF(hipGetLastError());
      }
      GAS_RETURN_IF(hipMemcpyAsync(ws.h_sizes, sz_out, sizeof(int32_t), hipMemcpyDeviceToHost, s));
      GAS_RETURN_IF(hipMemcpyAsync(ws.h_sizes + 2, ws.counters, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
      GAS_RETURN_IF(hipStreamSynchronize(s));
      const int32_t next = ws.h_sizes[0];
      unsigned long long seen;
      std::memcpy(&seen, ws.h_sizes + 2, sizeof(seen));
      history.push_back({nf, (int64_t)seen - total_edges, next, lbs ? 1 : 0});
      total_edges = (int64_t)seen;
      if (verbose) std::printf("iteration %d: frontier %d -> %d (%s)\n", it, nf, next, lbs ? "lbs" : "tiered");
      nf = next;
    }
    return hipSuccess;
  }

 private:
  void tb(const char *name) {
    if (timer) timer->begin(name);
  }
  void te(const char *name) {
    if (timer) timer->end(name);
  }
};

}  // namespace GASengine
