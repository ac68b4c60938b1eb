// Adapters from a vertex Program (the reference's functor concept,
// doc/mapgraph-overview.h:246-255; e.g. Algorithms/SSSP/sssp.h:34-393) to the
// engine's kernels: the expand/contract edge op used by the LBS and tiered
// expand kernels (the gather's item op is in gather.h).
//
// Functor arities: the canonical forms are those of sssp.h / cc.h /
// PageRank.h (apply: 6 args, contract: 6 args).  The fork's BFS contract
// (Algorithms/BFS/bfs.h:252, 8 args with rank_id and the visited bitmap) and
// the 7-arg apply (Algorithms/Swarm_similarity/bfs.h:198, rank_id first) are
// accepted too, detected at compile time.
#pragma once

#include <type_traits>

#include "Enums.h"
#include "util.h"

namespace GASengine {

// Non-owning view of a resident graph (CSR + CSC).  For undirected graphs the
// CSC aliases the CSR and csc_edge_ids is NULL (slot k of v's in-list is then
// the CSR entry v->u, whose value equals u->v for the symmetric builders).
struct DeviceGraph {
  int32_t nodes = 0;
  int64_t edges = 0;
  int directed = 0;
  const int32_t *row_offsets = nullptr;
  const int32_t *column_indices = nullptr;
  const int32_t *column_offsets = nullptr;
  const int32_t *row_indices = nullptr;
  const int32_t *csc_edge_ids = nullptr;
  void *edge_values = nullptr;
};

namespace detail {

template <class P>
using GT = typename P::GatherType;
template <class P>
using MT = typename P::MiscType;
template <class P>
using VT = typename P::VertexType;
template <class P>
using ET = typename P::EdgeType;

// Parameter count of a functor's operator() (read from its type only, so it
// works for __device__ members in host and device passes alike).
template <class M>
struct arity_of;
template <class R, class C, class... A>
struct arity_of<R (C::*)(A...)> {
  static constexpr int value = sizeof...(A);
};
template <class R, class C, class... A>
struct arity_of<R (C::*)(A...) const> {
  static constexpr int value = sizeof...(A);
};
template <class F>
constexpr int arity() {
  return arity_of<decltype(&F::operator())>::value;
}

template <class P>
__device__ __forceinline__ void call_apply(int v, int iteration, GT<P> g, VT<P> &vl, ET<P> &el, char &changed) {
  typename P::apply f;
  if constexpr (arity<typename P::apply>() == 6)
    f(v, iteration, g, vl, el, changed);
  else
    f(0, v, iteration, g, vl, el, changed);  // 7-arg form: (rank_id, vertex, iteration, ...)
}

template <class P>
__device__ __forceinline__ void call_contract(int iteration, int &v, VT<P> &vl, ET<P> &el, GT<P> *tmp, MT<P> &misc,
                                              unsigned char *bitmap) {
  typename P::contract f;
  if constexpr (arity<typename P::contract>() == 6)
    f(iteration, v, vl, el, tmp, misc);
  else
    f(0, iteration, bitmap, v, vl, el, tmp, misc);  // 8-arg form: (rank_id, iteration, bitmap, ...)
}

template <class P, class = void>
struct has_gather_vertex : std::false_type {};
template <class P>
struct has_gather_vertex<P, std::void_t<typename P::gather_vertex>> : std::true_type {};

}  // namespace detail

// k_gas_order: workgroups of kOrderBlock threads, at most kOrderGridMax of
// them (each thread a run of consecutive bitmap words), so that a
// workgroup's threads read all its predecessors' counts at once.
constexpr int kOrderBlock = 512;
constexpr int kOrderGridMax = kOrderBlock;

// Device-side loop state of a GAS run.  k_gas_ctl, the first kernel of every
// iteration, decides on the device what the host decides in the reference
// (enactor_vertex_centric.cuh:3637-3667, 3801-3804): stop when the frontier
// is empty or iter_num iterations have run, else choose the expansion
// (frontier > threshold: LBS, otherwise the tiered kernel).  The kernels of
// the iteration read the frontier size from the field meant for them
// (lbs_size, tier_size, gather_size), which is 0 for the path not taken and
// for every kernel once the loop is done, so a captured sequence of
// iterations runs exactly the iterations the run needs.
struct GasCtl {
  int32_t iter;        // number of the iteration being run (what the functors see)
  int32_t done;
  int32_t lbs_size, tier_size, gather_size;
  int32_t threshold, iter_num;
  int32_t iterations;  // iterations run so far
  int32_t lbs_iters;   // of which expanded by LBS
  int32_t order_epoch; // k_gas_order runs so far (its parity picks the scan words)
  int32_t tier_max_degree;  // > 0: a frontier within the threshold still goes to LBS when one
                            // of its rows is longer than this (0: the threshold alone decides)
  long long sum_frontier, sum_next;
};

// Iteration state shared by the kernels of one GAS iteration (by value).
template <class Program>
struct GasState {
  DeviceGraph g;
  typename Program::VertexType vl;
  typename Program::EdgeType el;
  typename Program::GatherType *gather_tmp;
  char *changed;
  uint32_t *dedup_next;  // bitmap deduplicating the next frontier
  uint32_t *dedup_prev;  // bitmap that deduplicated the current frontier (cleared here)
  int iteration;
  int in_edges;  // 0: expand over CSR out-edges, 1: over CSC in-edges
  int do_apply;
  int do_post;
  const int32_t *iter_ptr;  // device-driven loops: the iteration number lives here (GasCtl::iter)
  int32_t *preds;           // MARK_PREDECESSORS: discovering vertex per vertex, else NULL
  int append;               // 0: the next frontier is only marked in dedup_next (k_gas_order lists it)

  __device__ __forceinline__ int iter() const { return iter_ptr ? *iter_ptr : iteration; }
};

// Expand + contract fused per edge: expand_edge -> contract -> bitmap dedup
// -> enqueue.  The reference materialises the edge frontier (keys + misc,
// 8 B/edge) between its expand and contract kernels; here the contract of an
// edge runs right after its expand, which keeps the reference's per-element
// order (contract on every emitted element, then the visited-bitmap cull,
// contract_atomic/cta.cuh:529-629).
template <class Program>
struct ProgramOp {
  static constexpr bool kItemValues = false;
  GasState<Program> s;

  __device__ __forceinline__ void prologue(int32_t v) {
    s.dedup_prev[v >> 5] = 0u;  // every set bit there belongs to a current-frontier vertex
    if (s.do_apply) {
      char ch = s.changed[v];
      detail::call_apply<Program>(v, s.iter(), s.gather_tmp[v], s.vl, s.el, ch);
      s.changed[v] = ch;
    }
    if (s.do_post) {
      typename Program::post_apply f;
      f(v, s.vl, s.el, s.gather_tmp);
    }
  }

  __device__ __forceinline__ int32_t degree(int32_t v) {
    typename Program::expand_vertex f;
    if (!f(v, s.changed[v], s.vl, s.el)) return 0;
    const int32_t *off = s.in_edges ? s.g.column_offsets : s.g.row_offsets;
    return off[v + 1] - off[v];
  }

  __device__ __forceinline__ int32_t row_begin(int32_t v) {
    return (s.in_edges ? s.g.column_offsets : s.g.row_offsets)[v];
  }

  __device__ __forceinline__ int32_t item_value(int32_t) { return 0; }

  // Staged form for k_lbs_expand: the neighbour ids of a thread's edges load
  // together, then each edge runs the functors (expand_edge, contract) once.
  struct Loads {
    int32_t nb, eid;
  };
  __device__ __forceinline__ Loads load(int32_t, int32_t, int32_t slot) {
    if (s.in_edges) return {s.g.row_indices[slot], s.g.csc_edge_ids ? s.g.csc_edge_ids[slot] : slot};
    return {s.g.column_indices[slot], slot};
  }

  __device__ __forceinline__ bool apply(int32_t v, const Loads &l, int32_t &out) {
    int frontier = -1;
    typename Program::MiscType misc = typename Program::MiscType();
    typename Program::expand_edge fe;
    const int it = s.iter();
    fe(true, it, v, l.nb, l.eid, s.vl, s.el, frontier, misc);
    if (frontier == -1) return false;
    detail::call_contract<Program>(it, frontier, s.vl, s.el, s.gather_tmp, misc,
                                   reinterpret_cast<unsigned char *>(s.dedup_next));
    if (frontier == -1) return false;
    if (!s.append && !s.preds) {
      // Only marked (k_gas_order lists the next frontier from the bitmap):
      // the caller sets the bit (emit), a whole wave's marks together.
      out = frontier;
      return true;
    }
    if (!bit_test_and_set(s.dedup_next, frontier)) return false;
    if (s.preds) s.preds[frontier] = v;  // the winner of the dedup bit discovered it
    if (!s.append) return false;
    out = frontier;
    return true;
  }

  // What k_lbs_expand does with apply()'s result, on every lane of the wave.
  __device__ __forceinline__ void emit(bool keep, int32_t out, int32_t *s_out, int32_t *s_nout) {
    if (!s.append && !s.preds)
      wave_mark(keep, out, s.dedup_next);
    else
      wave_append(keep, out, s_out, s_nout);
  }

  __device__ __forceinline__ bool edge(int32_t v, int32_t val, int32_t slot, int32_t &out) {
    int32_t o = 0;
    const bool keep = apply(v, load(v, val, slot), o);
    if (keep && !s.append && !s.preds) {
      bit_test_and_set(s.dedup_next, o);
      return false;
    }
    out = o;
    return keep;
  }

  __device__ __forceinline__ void finish() {}
};

}  // namespace GASengine
