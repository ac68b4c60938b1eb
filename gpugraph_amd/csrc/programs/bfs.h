// BFS vertex program (reference: Algorithms/BFS/bfs.h:28-298).
// label[src] = 0, every other label = -1 (INIT_VALUE); every out-neighbour is
// emitted; contract keeps a vertex the first time it is reached and stamps it
// with iteration + 1, so labels are BFS depths.
#pragma once

#include "../gasengine/csr_problem.h"

struct bfs {
  typedef int DataType;
  typedef DataType MiscType;
  typedef DataType GatherType;
  typedef int VertexId;
  typedef int SizeT;

  static const DataType INIT_VALUE = -1;
  static const bool allow_duplicates = true;

  struct VertexType {
    int *d_labels;
    int *d_labels_src;
    int nodes;
    int edges;
    VertexType() : d_labels(nullptr), d_labels_src(nullptr), nodes(0), edges(0) {}
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
    (void)directed; (void)d_row_offsets; (void)d_column_indices; (void)d_column_offsets;
    (void)d_row_indices; (void)d_edge_values; (void)d_frontier_values; (void)edge_list;
    hipStream_t s = GASengine::engine_stream();
    vertex_list.nodes = nodes;
    vertex_list.edges = edges;
    if (!vertex_list.d_labels) (void)hipMalloc(&vertex_list.d_labels, sizeof(int) * (size_t)(nodes > 0 ? nodes : 1));
    vertex_list.d_labels_src = vertex_list.d_labels;
    (void)GASengine::fill_async<int>(vertex_list.d_labels, INIT_VALUE, nodes, s);
    for (int i = 0; i < num_srcs; i++) (void)GASengine::fill_async<int>(vertex_list.d_labels + srcs[i], 0, 1, s);
    (void)hipMemcpyAsync(d_frontier_keys[0], srcs, sizeof(int) * (size_t)num_srcs, hipMemcpyHostToDevice, s);
    (void)hipStreamSynchronize(s);
  }

  static void Release(VertexType &vertex_list, EdgeType &) {
    if (vertex_list.d_labels) (void)hipFree(vertex_list.d_labels);
    vertex_list.d_labels = vertex_list.d_labels_src = nullptr;
  }

  static SrcVertex srcVertex() { return SINGLE; }
  static GatherEdges gatherOverEdges() { return NO_GATHER_EDGES; }
  static ApplyVertices applyOverEdges() { return NO_APPLY_VERTICES; }
  static ExpandEdges expandOverEdges() { return EXPAND_OUT_EDGES; }
  static PostApplyVertices postApplyOverEdges() { return NO_POST_APPLY_VERTICES; }

  struct gather_sum {
    __device__ GatherType operator()(GatherType left, GatherType right) { return max(left, right); }
  };

  struct gather_vertex {
    __device__ void operator()(const int, const GatherType, VertexType &, EdgeType &) {}
  };

  // Pull form: a neighbour one level above is a parent (bfs.h gather_edge).
  struct gather_edge {
    __device__ void operator()(const int vertex_id, const int, const int neighbor_id_in, VertexType &vertex_list,
                               EdgeType &, GatherType &new_value) {
      int nb_label = vertex_list.d_labels_src[neighbor_id_in];
      int my_label = vertex_list.d_labels[vertex_id];
      new_value = (my_label - nb_label) == 1 ? neighbor_id_in : -1;
    }
  };

  struct apply {
    __device__ void operator()(const int, const int, GatherType, VertexType &, EdgeType &, char &changed) {
      changed = 1;
    }
  };

  struct post_apply {
    __device__ void operator()(const int, VertexType &, EdgeType &, GatherType *) {}
  };

  struct expand_vertex {
    __device__ bool operator()(const int, const char, VertexType &, EdgeType &) { return true; }
  };

  struct expand_edge {
    __device__ void operator()(const bool, const int, const int, const int neighbor_id_in, const int, VertexType &,
                               EdgeType &, int &frontier, int &) {
      frontier = neighbor_id_in;
    }
  };

  // Keep the first visit only; stamp the depth.  Concurrent first visits in
  // one iteration all write the same value; the engine's dedup bitmap then
  // enqueues the vertex once.
  struct contract {
    __device__ void operator()(const int iteration, int &vertex_id, VertexType &vertex_list, EdgeType &,
                               GatherType *, int &) {
      if (vertex_list.d_labels[vertex_id] != -1)
        vertex_id = -1;
      else
        vertex_list.d_labels[vertex_id] = iteration + 1;
    }
  };

  static void extractResult(VertexType &vertex_list, DataType *h_output) {
    (void)hipMemcpy(h_output, vertex_list.d_labels, sizeof(DataType) * (size_t)vertex_list.nodes,
                    hipMemcpyDeviceToHost);
  }

  // Device pointer of the result array (used by the C ABI for device outputs).
  static DataType *resultPointer(VertexType &vertex_list) { return vertex_list.d_labels; }
};
