// Single-source shortest paths (reference: Algorithms/SSSP/sssp.h:34-393).
// dist = 100000000 except the sources (0).  expand_edge emits a neighbour
// whose distance would drop (dst > src + w) with misc = src + w; contract
// folds misc into gather_tmp with atomicMin; the next iteration's apply takes
// min(dist, gathered) and flags the vertex changed; post_apply commits and
// resets gather_tmp.  Int32 weights from the graph's edge values.
#pragma once

#include "../gasengine/csr_problem.h"

struct sssp {
  typedef int DataType;
  typedef DataType MiscType;
  typedef DataType GatherType;
  typedef int VertexId;
  typedef int SizeT;

  static const int INIT_VALUE = 100000000;
  static const bool allow_duplicates = true;

  struct VertexType {
    int nodes;
    int edges;
    DataType *d_dists_out;
    DataType *d_dists;
    VertexType() : nodes(0), edges(0), d_dists_out(nullptr), d_dists(nullptr) {}
  };

  struct EdgeType {
    int nodes;
    int edges;
    DataType *d_weights;
    EdgeType() : nodes(0), edges(0), d_weights(nullptr) {}
  };

  static void Initialize(const int directed, const int nodes, const int edges, int num_srcs, int *srcs,
                         int *d_row_offsets, int *d_column_indices, int *d_column_offsets, int *d_row_indices,
                         DataType *d_edge_values, VertexType &vertex_list, EdgeType &edge_list,
                         int *d_frontier_keys[3], MiscType *d_frontier_values[3]) {
    (void)directed; (void)d_row_offsets; (void)d_column_indices; (void)d_column_offsets;
    (void)d_row_indices; (void)d_frontier_values;
    hipStream_t s = GASengine::engine_stream();
    vertex_list.nodes = nodes;
    vertex_list.edges = edges;
    edge_list.nodes = nodes;
    edge_list.edges = edges;
    const size_t nn = (size_t)(nodes > 0 ? nodes : 1);
    if (!vertex_list.d_dists) (void)hipMalloc(&vertex_list.d_dists, sizeof(DataType) * nn);
    if (!vertex_list.d_dists_out) (void)hipMalloc(&vertex_list.d_dists_out, sizeof(DataType) * nn);
    // The weights stay in the graph's resident value array (the reference
    // copies them into a second device buffer, sssp.h:103-114).
    edge_list.d_weights = d_edge_values;
    (void)GASengine::fill_async<int>(vertex_list.d_dists, INIT_VALUE, nodes, s);
    for (int i = 0; i < num_srcs; i++) (void)GASengine::fill_async<int>(vertex_list.d_dists + srcs[i], 0, 1, s);
    (void)hipMemcpyAsync(vertex_list.d_dists_out, vertex_list.d_dists, sizeof(DataType) * (size_t)nodes,
                         hipMemcpyDeviceToDevice, s);
    (void)hipMemcpyAsync(d_frontier_keys[0], srcs, sizeof(int) * (size_t)num_srcs, hipMemcpyHostToDevice, s);
    (void)hipStreamSynchronize(s);
  }

  static void Release(VertexType &vertex_list, EdgeType &) {
    if (vertex_list.d_dists) (void)hipFree(vertex_list.d_dists);
    if (vertex_list.d_dists_out) (void)hipFree(vertex_list.d_dists_out);
    vertex_list.d_dists = vertex_list.d_dists_out = nullptr;
  }

  static SrcVertex srcVertex() { return SINGLE; }
  static GatherEdges gatherOverEdges() { return NO_GATHER_EDGES; }
  static ApplyVertices applyOverEdges() { return APPLY_FRONTIER; }
  static PostApplyVertices postApplyOverEdges() { return POST_APPLY_FRONTIER; }
  static ExpandEdges expandOverEdges() { return EXPAND_OUT_EDGES; }

  struct gather_vertex {
    __device__ void operator()(const int, const GatherType, VertexType &, EdgeType &) {}
  };

  struct gather_edge {
    __device__ void operator()(const int, const int, const int, VertexType &, EdgeType &, GatherType &) {}
  };

  struct gather_sum {
    __device__ GatherType operator()(const GatherType &left, const GatherType &right) { return left + right; }
  };

  struct apply {
    __device__ void operator()(const int vertex_id, const int iteration, GatherType gathervalue,
                               VertexType &vertex_list, EdgeType &, char &changed) {
      const int oldvalue = vertex_list.d_dists[vertex_id];
      const int newvalue = min(oldvalue, gathervalue);
      changed = (iteration == 0 || oldvalue != newvalue) ? 1 : 0;
      vertex_list.d_dists_out[vertex_id] = newvalue;
    }
  };

  struct post_apply {
    __device__ void operator()(const int vertex_id, VertexType &vertex_list, EdgeType &, GatherType *gather_tmp) {
      vertex_list.d_dists[vertex_id] = vertex_list.d_dists_out[vertex_id];
      gather_tmp[vertex_id] = INIT_VALUE;
    }
  };

  struct expand_vertex {
    __device__ bool operator()(const int, const char changed, VertexType &, EdgeType &) { return changed == 1; }
  };

  struct expand_edge {
    __device__ void operator()(const bool, const int, const int vertex_id, const int neighbor_id_in,
                               const int edge_id, VertexType &vertex_list, EdgeType &edge_list, int &frontier,
                               int &misc_value) {
      const int src_dist = vertex_list.d_dists[vertex_id];
      const int dst_dist = vertex_list.d_dists[neighbor_id_in];
      const DataType w = edge_list.d_weights[edge_id];
      frontier = dst_dist > src_dist + w ? neighbor_id_in : -1;
      misc_value = src_dist + w;
    }
  };

  // atomicMin as the reference's contract (sssp.h:383), after a plain read:
  // gather_tmp only decreases during an iteration, so a read that is not
  // larger than misc_value means the atomic could not change it.
  struct contract {
    __device__ void operator()(const int, int &vertex_id, VertexType &, EdgeType &, GatherType *gather_tmp,
                               int &misc_value) {
      if (misc_value < gather_tmp[vertex_id]) atomicMin(&gather_tmp[vertex_id], misc_value);
    }
  };

  static void extractResult(VertexType &vertex_list, DataType *h_output) {
    (void)hipMemcpy(h_output, vertex_list.d_dists, sizeof(DataType) * (size_t)vertex_list.nodes,
                    hipMemcpyDeviceToHost);
  }

  static DataType *resultPointer(VertexType &vertex_list) { return vertex_list.d_dists; }
};
