// Vertex-program variants that traverse in-edges or both edge directions.
//
// The reference engine accepts EXPAND_IN_EDGES / EXPAND_ALL_EDGES and
// GATHER_OUT_EDGES / GATHER_ALL_EDGES (GASengine/Enums.h, handled in
// enactor_vertex_centric.cuh:893-1048, 2442-2711), but none of its shipped
// programs asks for them.  These variants of the shipped programs do, on a
// directed graph:
//   bfs_reverse   expand over in-edges: depth = hops from each vertex to src
//   cc_weak       expand over all edges: labels of the weakly connected
//                 components (min vertex id)
//   pagerank_all  gather over all edges with degree = out + in, expand over
//                 all edges: the reference PageRank of the graph read as
//                 undirected (its undirected branch, PageRank.h:155-161,
//                 adds the in-degree to the out-degree)
// Each differs from its base program only in the direction functions (and
// pagerank_all in the degree it divides by).
#pragma once

#include "bfs.h"
#include "cc.h"
#include "pagerank.h"

struct bfs_reverse : bfs {
  static ExpandEdges expandOverEdges() { return EXPAND_IN_EDGES; }
};

struct cc_weak : cc {
  static ExpandEdges expandOverEdges() { return EXPAND_ALL_EDGES; }
};

namespace pagerank_detail {
static __global__ void k_add_in_degree(const int *__restrict__ co, int *__restrict__ deg, int n) {
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) deg[v] += co[v + 1] - co[v];
}
}  // namespace pagerank_detail

struct pagerank_all : pagerank {
  static void Initialize(const int directed, const int nodes, const int edges, int num_srcs, int *srcs,
                         int *d_row_offsets, int *d_column_indices, int *d_column_offsets, int *d_row_indices,
                         DataType *d_edge_values, VertexType &vertex_list, EdgeType &edge_list,
                         int *d_frontier_keys[3], MiscType *d_frontier_values[3]) {
    pagerank::Initialize(directed, nodes, edges, num_srcs, srcs, d_row_offsets, d_column_indices, d_column_offsets,
                         d_row_indices, d_edge_values, vertex_list, edge_list, d_frontier_keys, d_frontier_values);
    if (nodes > 0)
      hipLaunchKernelGGL(pagerank_detail::k_add_in_degree, dim3(GASengine::grid_for(nodes, 256)), dim3(256), 0,
                         GASengine::engine_stream(), d_column_offsets, vertex_list.d_num_out_edge, nodes);
  }
  static GatherEdges gatherOverEdges() { return GATHER_ALL_EDGES; }
  static ExpandEdges expandOverEdges() { return EXPAND_ALL_EDGES; }
};
