// PageRank forward (FP32 Kahan gather, FP64 dangling sum and residual) and
// its adjoint.  Filled in below the first GPU milestone.
#include <hip/hip_runtime.h>

#include "internal.h"

extern "C" {

int gg_pagerank(gg_graph *g, float, const float *, double, int32_t, float *, uint32_t, gg_stats *) {
  gg::clear_error();
  GG_CHECK_ARG(g, "graph is NULL");
  return gg::set_error(GG_ERR_UNSUPPORTED, "gg_pagerank: not built yet");
}

int gg_pagerank_adjoint(gg_graph *g, float, const float *, const float *, double, int32_t, float *, float *, uint32_t,
                        gg_stats *) {
  gg::clear_error();
  GG_CHECK_ARG(g, "graph is NULL");
  return gg::set_error(GG_ERR_UNSUPPORTED, "gg_pagerank_adjoint: not built yet");
}

int gg_batch_run(gg_graph *g, const gg_batch_spec *, uint32_t, gg_stats *) {
  gg::clear_error();
  GG_CHECK_ARG(g, "graph is NULL");
  return gg::set_error(GG_ERR_UNSUPPORTED, "gg_batch_run: not built yet");
}

}  // extern "C"
