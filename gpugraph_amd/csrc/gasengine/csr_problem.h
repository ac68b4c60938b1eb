// CsrProblem: device graph + per-iteration state of a GAS run, the role of
// the reference's CsrProblem/GraphSlice (GASengine/csr_problem.cuh:117-1046).
//
// Layout in HBM (n vertices, m edges, all int32 unless noted):
//   CSR  row_offsets[n+1], column_indices[m], edge_values[m]
//   CSC  column_offsets[n+1], row_indices[m], csc_edge_ids[m]   (directed only;
//        undirected graphs use the CSR as their own CSC)
//   frontier[2][n]        ping-pong vertex queues (no O(m) edge frontier: the
//                         reference's EDGE_FRONTIERS queues are 1.5*m each)
//   dedup[2][ceil(n/32)]  bitmaps deduplicating the next frontier, alternated
//                         per iteration and cleared word-wise by the consumer
//   changed[n] (char), gather_tmp[n] (GatherType)
//   scan_local[n], partials[ceil(n/1024)], LbsInfo, sizes[2], counters[8]
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <optional>
#include <vector>

#include "Enums.h"
#include "config.h"
#include "csc_build.h"
#include "lbs.h"
#include "program_ops.h"

namespace GASengine {

// Stream on which programs' Initialize/extractResult issue their work; set by
// the enactor around those calls (the reference uses the default stream).
inline hipStream_t &engine_stream() {
  static thread_local hipStream_t s = nullptr;
  return s;
}

#define GAS_RETURN_IF(expr)            \
  do {                                 \
    hipError_t _e = (expr);            \
    if (_e != hipSuccess) return _e;   \
  } while (0)

// Engine buffers, sized for n vertices.  Reused across runs.
struct Workspace {
  int32_t capacity = 0;
  size_t gather_bytes = 0;
  int32_t *frontier[2] = {nullptr, nullptr};
  uint32_t *dedup[2] = {nullptr, nullptr};
  char *changed = nullptr;
  void *gather_tmp = nullptr;
  int32_t *scan_local = nullptr;
  int32_t *partials = nullptr;
  int32_t *parts = nullptr;  // LBS tile boundaries
  LbsInfo *info = nullptr;
  int32_t *sizes = nullptr;                   // [0],[1]: frontier sizes
  unsigned long long *counters = nullptr;     // [0]: edges inspected
  int32_t *frontier_values[3] = {nullptr, nullptr, nullptr};
  int32_t *preds = nullptr;                   // MARK_PREDECESSORS: discovering vertex per vertex
  int32_t *h_sizes = nullptr;                 // pinned readback: [0] size, [2..3] edge counter
  GasCtl *ctl = nullptr;                      // device loop state
  GasCtl *h_ctl = nullptr;                    // pinned read-back of ctl, sizes and the edge counter
  void *gcarry = nullptr;                     // gather tile carries
  size_t gcarry_bytes = 0;
  uint64_t *order_words = nullptr;            // k_gas_order's scan words (two sets)
  // Device-driven loops: a captured sequence of iterations, the arguments it
  // was captured with, and the stream it is captured on (launched on the
  // problem's stream, which may be the null stream).
  hipStream_t cap_stream = nullptr;
  hipGraphExec_t loop_exec = nullptr;
  std::vector<uint64_t> loop_key;

  static int32_t words(int32_t n) { return (n + 31) / 32; }

  hipError_t Allocate(int32_t n, int64_t m, size_t gather_elem_bytes) {
    if (capacity >= n && gather_bytes >= gather_elem_bytes * (size_t)n && frontier[0]) return hipSuccess;
    Release();
    const size_t nn = (size_t)(n > 0 ? n : 1);
    const size_t nw = (size_t)words(n > 0 ? n : 1);
    const size_t np = (nn + kCountTile - 1) / kCountTile + 1;
    GAS_RETURN_IF(hipMalloc(&frontier[0], nn * 4));
    GAS_RETURN_IF(hipMalloc(&frontier[1], nn * 4));
    GAS_RETURN_IF(hipMalloc(&dedup[0], nw * 4));
    GAS_RETURN_IF(hipMalloc(&dedup[1], nw * 4));
    GAS_RETURN_IF(hipMalloc(&changed, nn));
    GAS_RETURN_IF(hipMalloc(&gather_tmp, nn * (gather_elem_bytes > 4 ? gather_elem_bytes : 4)));
    GAS_RETURN_IF(hipMalloc(&scan_local, nn * 4));
    GAS_RETURN_IF(hipMalloc(&partials, np * 4));
    GAS_RETURN_IF(hipMalloc(&parts, lbs_partition_entries(n, m) * 4));
    GAS_RETURN_IF(hipMalloc(&info, sizeof(LbsInfo)));
    GAS_RETURN_IF(hipMalloc(&sizes, 4 * sizeof(int32_t)));
    GAS_RETURN_IF(hipMalloc(&counters, 8 * sizeof(unsigned long long)));
    GAS_RETURN_IF(hipHostMalloc(&h_sizes, 8 * sizeof(int32_t)));
    GAS_RETURN_IF(hipMalloc(&ctl, sizeof(GasCtl)));
    GAS_RETURN_IF(hipMalloc(&order_words, 2 * kOrderGridMax * sizeof(uint64_t)));
    GAS_RETURN_IF(hipHostMalloc(&h_ctl, 2 * sizeof(GasCtl)));
    capacity = n;
    gather_bytes = gather_elem_bytes * nn;
    return hipSuccess;
  }

  hipError_t EnsureCarry(size_t bytes) {
    if (gcarry_bytes >= bytes && gcarry) return hipSuccess;
    if (gcarry) (void)hipFree(gcarry);
    gcarry = nullptr;
    gcarry_bytes = 0;
    GAS_RETURN_IF(hipMalloc(&gcarry, bytes));
    gcarry_bytes = bytes;
    return hipSuccess;
  }

  void ResetLoop() {
    if (loop_exec) (void)hipGraphExecDestroy(loop_exec);
    loop_exec = nullptr;
    loop_key.clear();
  }

  hipError_t AllocatePreds() {
    if (!preds) GAS_RETURN_IF(hipMalloc(&preds, (size_t)(capacity > 0 ? capacity : 1) * 4));
    return hipSuccess;
  }

  // Misc-value queues handed to Program::Initialize (d_frontier_values[3]).
  hipError_t AllocateValues() {
    for (int i = 0; i < 3; i++)
      if (!frontier_values[i]) GAS_RETURN_IF(hipMalloc(&frontier_values[i], (size_t)(capacity > 0 ? capacity : 1) * 8));
    return hipSuccess;
  }

  void Release() {
    for (int i = 0; i < 2; i++) {
      if (frontier[i]) (void)hipFree(frontier[i]);
      if (dedup[i]) (void)hipFree(dedup[i]);
      frontier[i] = nullptr;
      dedup[i] = nullptr;
    }
    for (int i = 0; i < 3; i++) {
      if (frontier_values[i]) (void)hipFree(frontier_values[i]);
      frontier_values[i] = nullptr;
    }
    ResetLoop();
    if (cap_stream) (void)hipStreamDestroy(cap_stream);
    cap_stream = nullptr;
    void *ps[] = {changed, gather_tmp, scan_local, partials, parts, info, sizes, counters, preds, ctl, gcarry,
                  order_words};
    for (void *p : ps)
      if (p) (void)hipFree(p);
    if (h_sizes) (void)hipHostFree(h_sizes);
    if (h_ctl) (void)hipHostFree(h_ctl);
    ctl = nullptr;
    h_ctl = nullptr;
    order_words = nullptr;
    gcarry = nullptr;
    gcarry_bytes = 0;
    changed = nullptr;
    gather_tmp = nullptr;
    scan_local = nullptr;
    partials = nullptr;
    parts = nullptr;
    info = nullptr;
    sizes = nullptr;
    counters = nullptr;
    preds = nullptr;
    h_sizes = nullptr;
    capacity = 0;
    gather_bytes = 0;
  }
};

namespace detail {
template <class P, class = void>
struct has_extract_pred : std::false_type {};
template <class P>
struct has_extract_pred<P, std::void_t<decltype(&P::extractPred)>> : std::true_type {};
}  // namespace detail

// MARK_PREDECESSORS: the engine also records, for every vertex that enters a
// frontier, the vertex whose edge put it there first (the one that won the
// frontier's dedup bit); sources are their own predecessor, unreached
// vertices -1.  ExtractPreds copies them out (through Program::extractPred
// when the program has one, as csr_problem.cuh:459-483 does).  In BFS this
// is a BFS tree: each predecessor sits one level up.
template <class Program, typename VertexId = int, typename SizeT = int,
          typename Value = typename Program::DataType, bool MARK_PREDECESSORS = false, bool WITH_VALUE = true>
struct CsrProblem {
  typedef Program ProgramType;
  typedef typename Program::DataType DataType;
  static constexpr bool kMarkPredecessors = MARK_PREDECESSORS;

  DeviceGraph graph;
  Workspace ws;
  typename Program::VertexType vertex_list;
  typename Program::EdgeType edge_list;
  hipStream_t stream = nullptr;
  std::vector<void *> owned;  // device copies made by FromHostProblem
  // The driver's parameters when built from a Config (held, not default
  // constructed: a Config's first construction runs the driver's
  // registerParameters(), which only programs that use Config define).
  std::optional<Config> cfg;

  CsrProblem() = default;
  // Reference constructor (csr_problem.cuh:378): keeps the driver's
  // parameters; the graph comes with FromHostProblem.
  explicit CsrProblem(const Config &cfg_) : cfg(cfg_) {}
  CsrProblem(const CsrProblem &) = delete;
  CsrProblem &operator=(const CsrProblem &) = delete;
  ~CsrProblem() {
    ws.Release();
    for (void *p : owned) (void)hipFree(p);
  }

  // Share a graph already resident on the device (no copy).
  hipError_t Attach(const DeviceGraph &g, hipStream_t s) {
    graph = g;
    stream = s;
    GAS_RETURN_IF(ws.Allocate(g.nodes, g.edges, sizeof(typename Program::GatherType)));
    if (MARK_PREDECESSORS) GAS_RETURN_IF(ws.AllocatePreds());
    return hipSuccess;
  }

  // Reference form with the device and MPI rank (csr_problem.cuh:512-524):
  // the graph goes to device `device_id`; rank_id has no use on one GPU.
  hipError_t FromHostProblem(bool stream_from_host, SizeT nodes, SizeT edges, VertexId *h_column_indices,
                             SizeT *h_row_offsets, Value *h_edge_values, VertexId *h_row_indices,
                             SizeT *h_column_offsets, int num_gpus, int directed, int device_id, int /*rank_id*/) {
    GAS_RETURN_IF(hipSetDevice(device_id));
    return FromHostProblem(stream_from_host, nodes, edges, h_column_indices, h_row_offsets, h_edge_values,
                           h_row_indices, h_column_offsets, num_gpus, directed);
  }

  // Reference form (csr_problem.cuh:512-524, upstream 10-argument overload):
  // copy a host CSR to the current device; for directed graphs the CSC and
  // its edge-id permutation are built there (device_csr_to_csc).  A CSC the
  // caller passes is not needed and not read.
  hipError_t FromHostProblem(bool /*stream_from_host*/, SizeT nodes, SizeT edges, VertexId *h_column_indices,
                             SizeT *h_row_offsets, Value *h_edge_values, VertexId *h_row_indices,
                             SizeT *h_column_offsets, int /*num_gpus*/, int directed) {
    DeviceGraph g;
    g.nodes = nodes;
    g.edges = edges;
    g.directed = directed;
    auto up = [&](const void *src, size_t bytes, void **dst) -> hipError_t {
      GAS_RETURN_IF(hipMalloc(dst, bytes > 0 ? bytes : 4));
      owned.push_back(*dst);
      if (bytes) GAS_RETURN_IF(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
      return hipSuccess;
    };
    void *p = nullptr;
    GAS_RETURN_IF(up(h_row_offsets, sizeof(SizeT) * ((size_t)nodes + 1), &p));
    g.row_offsets = (const int32_t *)p;
    GAS_RETURN_IF(up(h_column_indices, sizeof(VertexId) * (size_t)edges, &p));
    g.column_indices = (const int32_t *)p;
    if (WITH_VALUE && h_edge_values) {
      GAS_RETURN_IF(up(h_edge_values, sizeof(Value) * (size_t)edges, &p));
      g.edge_values = p;
    }
    (void)h_row_indices;
    (void)h_column_offsets;
    if (directed) {
      auto alloc = [&](size_t bytes, void **dst) -> hipError_t {
        GAS_RETURN_IF(hipMalloc(dst, bytes > 0 ? bytes : 4));
        owned.push_back(*dst);
        return hipSuccess;
      };
      void *co = nullptr, *ri = nullptr, *eid = nullptr;
      GAS_RETURN_IF(alloc(4 * ((size_t)nodes + 1), &co));
      GAS_RETURN_IF(alloc(4 * (size_t)edges, &ri));
      GAS_RETURN_IF(alloc(4 * (size_t)edges, &eid));
      GAS_RETURN_IF(device_csr_to_csc(nodes, edges, g.row_offsets, g.column_indices, (int32_t *)co, (int32_t *)ri,
                                      (int32_t *)eid, stream));
      g.column_offsets = (const int32_t *)co;
      g.row_indices = (const int32_t *)ri;
      g.csc_edge_ids = (const int32_t *)eid;
    } else {
      g.column_offsets = g.row_offsets;
      g.row_indices = g.column_indices;
    }
    return Attach(g, stream);
  }

  // Clear per-run engine state (csr_problem.cuh:845-1045).  The frontier
  // contents come from Program::Initialize, called by the enactor.
  hipError_t Reset(FrontierType /*frontier_type*/ = VERTEX_FRONTIERS, double /*queue_sizing*/ = 1.0) {
    const int32_t n = graph.nodes;
    using G = typename Program::GatherType;
    GAS_RETURN_IF(fill_async<G>((G *)ws.gather_tmp, (G)Program::INIT_VALUE, n, stream));
    GAS_RETURN_IF(fill_async<char>(ws.changed, (char)0, n, stream));
    GAS_RETURN_IF(fill_async<uint32_t>(ws.dedup[0], 0u, Workspace::words(n), stream));
    GAS_RETURN_IF(fill_async<uint32_t>(ws.dedup[1], 0u, Workspace::words(n), stream));
    GAS_RETURN_IF(fill_async<unsigned long long>(ws.counters, 0ull, 8, stream));
    return hipSuccess;
  }

  // Copy the program's result to host memory (csr_problem.cuh:433-458).
  hipError_t ExtractResults(DataType *h_out) {
    GAS_RETURN_IF(hipStreamSynchronize(stream));
    engine_stream() = stream;
    Program::extractResult(vertex_list, h_out);
    return hipStreamSynchronize(stream);
  }

  // Predecessors of the last run (csr_problem.cuh:459-483); needs
  // MARK_PREDECESSORS.
  hipError_t ExtractPreds(int *h_preds) {
    if (!MARK_PREDECESSORS || !ws.preds) return hipErrorInvalidValue;
    GAS_RETURN_IF(hipStreamSynchronize(stream));
    if constexpr (detail::has_extract_pred<Program>::value) {
      engine_stream() = stream;
      Program::extractPred(ws.preds, graph.nodes, h_preds);
    } else {
      GAS_RETURN_IF(hipMemcpy(h_preds, ws.preds, sizeof(int) * (size_t)graph.nodes, hipMemcpyDeviceToHost));
    }
    return hipStreamSynchronize(stream);
  }
};

}  // namespace GASengine
