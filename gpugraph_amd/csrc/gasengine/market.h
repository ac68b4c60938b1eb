// Host-side graph types the reference's drivers use: CsrGraph
// (b40c/graph/csr_graph.cuh:69-370) and BuildMarketGraph
// (b40c/graph/builder/market.cuh:173-205), under the same namespaces so a
// driver's `using namespace b40c; using namespace graph;` and
// `builder::BuildMarketGraph<WITH_VALUES>(file, csr_graph, undirected)` keep
// compiling.  The file is parsed by the library's Matrix Market reader
// (gg_read_mtx: 1-based ids, rows sorted stably by (row, column), duplicates
// kept, mirrored when undirected; values read as VALUE_TYPE, 1 for pattern
// files); the arrays are host memory owned by the CsrGraph.  Links against
// libgasgraph.so.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>

#include "gasgraph.h"

namespace b40c {
namespace graph {

template <typename VertexId, typename Value, typename SizeT>
struct CsrGraph {
  SizeT nodes = 0;
  SizeT edges = 0;
  SizeT *row_offsets = nullptr;
  VertexId *column_indices = nullptr;
  SizeT *column_offsets = nullptr;  // not built: the engine builds its CSC on the device
  VertexId *row_indices = nullptr;
  Value *edge_values = nullptr;
  Value *node_values = nullptr;
  VertexId *from_nodes = nullptr;
  VertexId *to_nodes = nullptr;
  bool pinned = false;  // accepted for the reference's signature; storage is pageable here

  explicit CsrGraph(bool pinned_ = false) : pinned(pinned_) {}
  CsrGraph(const CsrGraph &) = delete;
  CsrGraph &operator=(const CsrGraph &) = delete;
  ~CsrGraph() { Free(); }

  void Free() {
    std::free(row_offsets);
    std::free(column_indices);
    std::free(column_offsets);
    std::free(row_indices);
    std::free(edge_values);
    std::free(node_values);
    std::free(from_nodes);
    std::free(to_nodes);
    row_offsets = nullptr;
    column_indices = nullptr;
    column_offsets = nullptr;
    row_indices = nullptr;
    edge_values = nullptr;
    node_values = nullptr;
    from_nodes = nullptr;
    to_nodes = nullptr;
    nodes = edges = 0;
  }
};

namespace builder {

// Returns 0 on success, -1 when the file cannot be read or parsed.
template <bool LOAD_VALUES, typename VertexId, typename Value, typename SizeT>
int BuildMarketGraph(char *mm_filename, CsrGraph<VertexId, Value, SizeT> &csr_graph, bool undirected) {
  static_assert(sizeof(VertexId) == 4 && sizeof(SizeT) == 4, "32-bit vertex ids and offsets");
  if (!mm_filename) {
    std::printf("Reading from stdin is not supported; pass a file\n");
    return -1;
  }
  const int32_t dtype =
      !LOAD_VALUES ? GG_VALUES_NONE : (std::is_floating_point<Value>::value ? GG_VALUES_F32 : GG_VALUES_I32);
  gg_csr c{};
  std::printf("Reading from %s:\n", mm_filename);
  if (gg_read_mtx(mm_filename, undirected ? 1 : 0, dtype, &c) != GG_OK) {
    std::printf("%s\n", gg_last_error());
    return -1;
  }
  csr_graph.Free();
  csr_graph.nodes = (SizeT)c.n;
  csr_graph.edges = (SizeT)c.m;
  const size_t ro_bytes = sizeof(SizeT) * ((size_t)c.n + 1);
  const size_t m = (size_t)c.m;
  csr_graph.row_offsets = (SizeT *)std::malloc(ro_bytes);
  csr_graph.column_indices = (VertexId *)std::malloc(sizeof(VertexId) * (m ? m : 1));
  std::memcpy(csr_graph.row_offsets, c.row_offsets, ro_bytes);
  if (m) std::memcpy(csr_graph.column_indices, c.col_indices, sizeof(VertexId) * m);
  if (LOAD_VALUES) {
    csr_graph.edge_values = (Value *)std::malloc(sizeof(Value) * (m ? m : 1));
    for (size_t e = 0; e < m; e++) {
      if (dtype == GG_VALUES_F32)
        csr_graph.edge_values[e] = (Value)static_cast<const float *>(c.values)[e];
      else
        csr_graph.edge_values[e] = (Value)static_cast<const int32_t *>(c.values)[e];
    }
  }
  gg_csr_free(&c);
  return 0;
}

}  // namespace builder
}  // namespace graph
}  // namespace b40c
