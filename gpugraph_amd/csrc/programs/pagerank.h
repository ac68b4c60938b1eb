// PageRank vertex program as the reference defines it (Algorithms/PageRank/
// PageRank.h:33-409): every vertex starts active with rank 0.15; gather over
// in-edges sums r_u / outdeg_u; apply sets r = 0.15 + 0.85 * sum and marks
// the vertex changed iff |delta| >= 0.01; a changed vertex activates its
// out-neighbours (deduplicated: allow_duplicates = false).  FP32.
//
// outdeg is the CSR out-degree, as in the CPU check (PageRank.cu:119-121).
// The reference's undirected branch (PageRank.h:155-161) adds CSC degrees,
// but its CsrProblem only allocates a CSC for directed graphs
// (csr_problem.cuh:607), so that branch has no defined result to match.
#pragma once

#include "../gasengine/csr_problem.h"

namespace pagerank_detail {
static __global__ void k_out_degree(const int *__restrict__ ro, int *__restrict__ deg, int n) {
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) deg[v] = ro[v + 1] - ro[v];
}
}  // namespace pagerank_detail

struct pagerank {
  typedef float DataType;
  typedef int MiscType;
  typedef float GatherType;
  typedef int VertexId;
  typedef int SizeT;

  static constexpr DataType INIT_VALUE = 0.0f;
  static const bool allow_duplicates = false;

  struct VertexType {
    int nodes;
    int edges;
    DataType *d_dists;
    int *d_num_out_edge;
    VertexType() : nodes(0), edges(0), d_dists(nullptr), d_num_out_edge(nullptr) {}
  };

  struct EdgeType {
    int nodes;
    int edges;
    EdgeType() : nodes(0), edges(0) {}
  };

  static void Initialize(const int directed, const int nodes, const int edges, int num_srcs, int *srcs,
                         int *d_row_offsets, int *d_column_indices, int *d_column_offsets, int *d_row_indices,
                         DataType *d_edge_values, VertexType &vertex_list, EdgeType &edge_list,
                         int *d_frontier_keys[3], MiscType *d_frontier_values[3]) {
    (void)directed; (void)num_srcs; (void)srcs; (void)d_column_indices; (void)d_column_offsets;
    (void)d_row_indices; (void)d_edge_values; (void)edge_list; (void)d_frontier_values;
    hipStream_t s = GASengine::engine_stream();
    vertex_list.nodes = nodes;
    vertex_list.edges = edges;
    const size_t nn = (size_t)(nodes > 0 ? nodes : 1);
    if (!vertex_list.d_dists) (void)hipMalloc(&vertex_list.d_dists, sizeof(DataType) * nn);
    if (!vertex_list.d_num_out_edge) (void)hipMalloc(&vertex_list.d_num_out_edge, sizeof(int) * nn);
    (void)GASengine::fill_async<float>(vertex_list.d_dists, 0.15f, nodes, s);
    (void)GASengine::sequence_async<int>(d_frontier_keys[0], nodes, s);
    if (nodes > 0)
      hipLaunchKernelGGL(pagerank_detail::k_out_degree, dim3(GASengine::grid_for(nodes, 256)), dim3(256), 0, s,
                         d_row_offsets, vertex_list.d_num_out_edge, nodes);
  }

  static void Release(VertexType &vertex_list, EdgeType &) {
    if (vertex_list.d_dists) (void)hipFree(vertex_list.d_dists);
    if (vertex_list.d_num_out_edge) (void)hipFree(vertex_list.d_num_out_edge);
    vertex_list.d_dists = nullptr;
    vertex_list.d_num_out_edge = nullptr;
  }

  static SrcVertex srcVertex() { return ALL; }
  static GatherEdges gatherOverEdges() { return GATHER_IN_EDGES; }
  static ApplyVertices applyOverEdges() { return APPLY_FRONTIER; }
  static PostApplyVertices postApplyOverEdges() { return POST_APPLY_FRONTIER; }
  static ExpandEdges expandOverEdges() { return EXPAND_OUT_EDGES; }

  struct gather_edge {
    __device__ void operator()(const int, const int, const int neighbor_id_in, VertexType &vertex_list, EdgeType &,
                               GatherType &new_value) {
      DataType nb_dist = vertex_list.d_dists[neighbor_id_in];
      new_value = nb_dist / (DataType)vertex_list.d_num_out_edge[neighbor_id_in];
    }
  };

  struct gather_sum {
    __device__ GatherType operator()(const GatherType &left, const GatherType &right) { return left + right; }
  };

  struct apply {
    __device__ void operator()(const int vertex_id, const int, GatherType gathervalue, VertexType &vertex_list,
                               EdgeType &, char &changed) {
      const DataType oldvalue = vertex_list.d_dists[vertex_id];
      const DataType newvalue = 0.15f + (1.0f - 0.15f) * gathervalue;
      changed = fabsf(oldvalue - newvalue) < 0.01f ? 0 : 1;
      vertex_list.d_dists[vertex_id] = newvalue;
    }
  };

  struct post_apply {
    __device__ void operator()(const int, VertexType &, EdgeType &, GatherType *) {}
  };

  struct expand_vertex {
    __device__ bool operator()(const int, const char changed, VertexType &, EdgeType &) { return changed; }
  };

  struct expand_edge {
    __device__ void operator()(const bool changed, const int, const int, const int neighbor_id_in, const int,
                               VertexType &, EdgeType &, int &frontier, int &) {
      if (changed) frontier = neighbor_id_in;
    }
  };

  struct contract {
    __device__ void operator()(const int, int &, VertexType &, EdgeType &, GatherType *, int &) {}
  };

  static void extractResult(VertexType &vertex_list, DataType *h_output) {
    (void)hipMemcpy(h_output, vertex_list.d_dists, sizeof(DataType) * (size_t)vertex_list.nodes,
                    hipMemcpyDeviceToHost);
  }

  static DataType *resultPointer(VertexType &vertex_list) { return vertex_list.d_dists; }
};
