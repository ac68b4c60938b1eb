// GAS configuration enums of the vertex-program API.  Same names and values
// as the reference (GASengine/Enums.h:1-40) so programs written against it
// compile unchanged.
#pragma once

enum SrcVertex { SINGLE, ALL };

enum GatherEdges { NO_GATHER_EDGES, GATHER_IN_EDGES, GATHER_OUT_EDGES, GATHER_ALL_EDGES };

enum ExpandEdges { NO_EXPAND_EDGES, EXPAND_IN_EDGES, EXPAND_OUT_EDGES, EXPAND_ALL_EDGES };

enum ApplyVertices { NO_APPLY_VERTICES, APPLY_ALL, APPLY_FRONTIER };

enum PostApplyVertices { NO_POST_APPLY_VERTICES, POST_APPLY_ALL, POST_APPLY_FRONTIER };

namespace GASengine {

// Frontier queue configurations (GASengine/Enums.h, csr_problem.cuh:69-108).
// This engine keeps vertex frontiers only (expand and contract are fused, so
// no O(m) edge frontier is materialised); the values are accepted by Reset
// for source compatibility.
enum FrontierType {
  VERTEX_FRONTIERS,
  EDGE_FRONTIERS,
  MIXED_FRONTIERS,
  MULTI_GPU_FRONTIERS,
  MULTI_GPU_VERTEX_FRONTIER_SCALE = 2,
};

}  // namespace GASengine
