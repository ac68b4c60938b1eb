// Load-balanced frontier gather (GATHER_IN_EDGES / GATHER_OUT_EDGES /
// GATHER_ALL_EDGES): gather_tmp[v] = gather_sum over the gather edges of v of
// gather_edge(v, e, nb), for every frontier vertex v, INIT_VALUE when v has
// none.  The reference materialises (destination, value) pairs for all
// gather edges and runs ReduceByKey (enactor_vertex_centric.cuh:2104-2424).
//
// Here the frontier's gather degrees go through the LBS count / scan /
// partition kernels (lbs.h), so the merge path of (frontier items, their
// gather edges) is cut into equal tiles whatever the degree skew, and per
// tile (k_lbs_gather):
//   1. the tile's item offsets, vertices and row starts are staged in LDS and
//      each thread finds the item of each of its VT path positions;
//   2. the edge values are computed with consecutive lanes on consecutive
//      edges into LDS;
//   3. thread t folds edges [VT t, VT t + VT) in order, one partial per run of
//      equal items;
//   4. a segmented inclusive scan of the threads' last-run partials (wave
//      shuffles, then the waves in order) gives each thread the partial of
//      its first item from the threads before it;
//   5. an item that starts and ends in the tile is finished at once; the item
//      still open at the tile's end and the one that began before the tile
//      leave their in-tile partials in the tile's GatherCarry, and
//      k_lbs_gather_fixup folds those in path order for every item that
//      spans tiles.
// Every fold happens in an order fixed by the frontier and the tile size, so
// results are the same run to run.  gather_sum must be associative (the
// reference's ReduceByKey assumes the same); GatherType is 4 or 8 bytes.
#pragma once

#include <type_traits>

#include "lbs.h"
#include "program_ops.h"

namespace GASengine {

constexpr int kGatherVt = kExpandVt;
constexpr int kGatherTile = kExpandBlock * kGatherVt;

// Item op for the LBS count kernel: degree = length of the gather row.
template <class Program>
struct GatherRows {
  static constexpr bool kItemValues = false;
  GasState<Program> s;
  int in_edges;  // 1: CSC in-edges, 0: CSR out-edges

  __device__ __forceinline__ const int32_t *offs() const { return in_edges ? s.g.column_offsets : s.g.row_offsets; }
  __device__ __forceinline__ void prologue(int32_t) {}
  __device__ __forceinline__ int32_t degree(int32_t v) { return offs()[v + 1] - offs()[v]; }
  __device__ __forceinline__ int32_t row_begin(int32_t v) { return offs()[v]; }
  __device__ __forceinline__ int32_t item_value(int32_t) { return 0; }
  __device__ __forceinline__ void finish() {}

  __device__ __forceinline__ typename Program::GatherType value(int32_t v, int32_t slot) {
    int32_t nb, eid;
    if (in_edges) {
      nb = s.g.row_indices[slot];
      eid = s.g.csc_edge_ids ? s.g.csc_edge_ids[slot] : slot;
    } else {
      nb = s.g.column_indices[slot];
      eid = slot;
    }
    typename Program::GatherType val;
    typename Program::gather_edge ge;
    ge(v, eid, nb, s.vl, s.el, val);
    return val;
  }

  // Final value of item v: INIT_VALUE without gather edges; folded into the
  // existing value on the second pass of GATHER_ALL_EDGES.
  template <bool ACCUM>
  __device__ __forceinline__ void done(int32_t v, bool has, typename Program::GatherType acc) {
    using G = typename Program::GatherType;
    typename Program::gather_sum gs;
    G r = has ? acc : (G)Program::INIT_VALUE;
    if (ACCUM) r = gs(s.gather_tmp[v], r);
    s.gather_tmp[v] = r;
    if constexpr (detail::has_gather_vertex<Program>::value) {
      typename Program::gather_vertex gv;
      gv(v, r, s.vl, s.el);
    }
  }
};

// A partial over a run of edges; has = false: no edge yet.
template <class G>
struct GPart {
  G val;
  int32_t has;
};

template <class G>
struct GatherCarry {
  G open, tail;
  int32_t open_i, tail_i;      // frontier item index, -1 if none
  int32_t open_has, tail_has;
};

template <class G>
__device__ __forceinline__ G shfl_up_any(G x, int o) {
  static_assert(sizeof(G) == 4 || sizeof(G) == 8, "GatherType of 4 or 8 bytes");
  if constexpr (sizeof(G) == 4) {
    int32_t b;
    __builtin_memcpy(&b, &x, 4);
    b = __shfl_up(b, o);
    __builtin_memcpy(&x, &b, 4);
  } else {
    long long b;
    __builtin_memcpy(&b, &x, 8);
    b = __shfl_up(b, o);
    __builtin_memcpy(&x, &b, 8);
  }
  return x;
}

// a then b.
template <class Program>
__device__ __forceinline__ GPart<typename Program::GatherType> gcat(GPart<typename Program::GatherType> a,
                                                                     GPart<typename Program::GatherType> b) {
  typename Program::gather_sum gs;
  if (a.has && b.has) return {gs(a.val, b.val), 1};
  return a.has ? a : b;
}

template <class Program, bool ACCUM>
__global__ void __launch_bounds__(kExpandBlock)
k_lbs_gather(GatherRows<Program> op, const int32_t *__restrict__ frontier, const int32_t *__restrict__ fsize,
             const int32_t *__restrict__ scan_local, const int32_t *__restrict__ partials,
             const LbsInfo *__restrict__ info, const int32_t *__restrict__ parts,
             GatherCarry<typename Program::GatherType> *__restrict__ carry) {
  using G = typename Program::GatherType;
  using P = GPart<G>;
  constexpr int TILE = kGatherTile, VT = kGatherVt, W = kExpandBlock / kWave;
  __shared__ int32_t s_offs[TILE + 2];  // s_offs[k]: offset of item i0 - 1 + k, k = 0..ni+1
  __shared__ int32_t s_vert[TILE + 1];
  __shared__ int32_t s_begin[TILE + 1];
  __shared__ int32_t s_slot[TILE];
  __shared__ G s_val[TILE];
  __shared__ P s_incl[kExpandBlock];
  __shared__ P s_wend[W];
  __shared__ int32_t s_whead[W];
  __shared__ GatherCarry<G> s_carry;
  const int32_t nf = *fsize;
  const int32_t total = info->edges;
  const int64_t path = info->path;
  const int tid = threadIdx.x, lane = lane_id(), wid = tid / kWave;
  typename Program::gather_sum gs;
  for (int64_t t = blockIdx.x; t * TILE < path; t += gridDim.x) {
    const int64_t d0 = t * TILE;
    const int64_t d1 = d0 + TILE < path ? d0 + TILE : path;
    const int32_t i0 = parts[t], i1 = parts[t + 1];
    const int32_t j0 = (int32_t)(d0 - i0), j1 = (int32_t)(d1 - i1);
    const int32_t ni = i1 - i0, ne = j1 - j0;
    for (int k = tid; k <= ni + 1; k += kExpandBlock) {
      const int32_t i = i0 - 1 + k;  // i <= i1 + 1 <= nf + 1
      s_offs[k] = i < 0 ? 0 : (i < nf ? lbs_offset(scan_local, partials, i) : total);
      if (k <= ni) {
        const int32_t v = (i >= 0 && i < nf) ? frontier[i] : -1;
        s_vert[k] = v;
        s_begin[k] = v >= 0 ? op.row_begin(v) : 0;
      }
    }
    if (tid == 0) {
      s_carry.open_i = -1;
      s_carry.tail_i = -1;
      s_carry.open_has = 0;
      s_carry.tail_has = 0;
    }
    __syncthreads();
    {
      const int32_t dl = tid * VT;
      if (dl < ni + ne) {
        int32_t lo = dl - ne > 0 ? dl - ne : 0, hi = dl < ni ? dl : ni;
        while (lo < hi) {
          const int32_t mid = (lo + hi) >> 1;
          if (s_offs[mid + 1] <= j0 + (dl - mid) - 1)
            lo = mid + 1;
          else
            hi = mid;
        }
        int32_t li = lo, lj = dl - lo;
#pragma unroll
        for (int k = 0; k < VT; k++) {
          if (li + lj >= ni + ne) break;
          if (li < ni && s_offs[li + 1] <= j0 + lj) {
            li++;
          } else {
            s_slot[lj] = li;
            lj++;
          }
        }
      }
    }
    __syncthreads();
    for (int32_t k = tid; k < ne; k += kExpandBlock) {
      const int32_t s = s_slot[k];
      s_val[k] = op.value(s_vert[s], s_begin[s] + (j0 + k - s_offs[s]));
    }
    __syncthreads();

    // Step 3: this thread's edges [e0, e1) in runs of equal items.
    const int32_t e0 = tid * VT, e1 = min(e0 + VT, ne);
    const bool has = e0 < ne;
    int32_t fs = -1, ls = -1;
    P first{G(), 0}, run{G(), 0};
    bool head = true;
    if (has) {
      fs = s_slot[e0];
      int32_t cs = fs;
      int nruns = 1;
      for (int32_t k = e0; k < e1; k++) {
        const int32_t sk = s_slot[k];
        if (sk != cs) {
          if (nruns == 1) {
            first = run;
          } else {
            // A run inside this thread: the item's whole in-tile sum; such an
            // item starts in the tile.
            if (s_offs[cs + 1] <= j1 && cs >= 1) op.template done<ACCUM>(s_vert[cs], run.has, run.val);
          }
          run = P{G(), 0};
          cs = sk;
          nruns++;
        }
        run = run.has ? P{gs(run.val, s_val[k]), 1} : P{s_val[k], 1};
      }
      ls = cs;
      head = nruns > 1 || e0 == 0 || s_slot[e0 - 1] != ls;
    }

    // Step 4: segmented inclusive scan of the last-run partials.
    P sc = run;
    int32_t f = head ? 1 : 0;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      P pp;
      pp.val = shfl_up_any(sc.val, o);
      pp.has = __shfl_up(sc.has, o);
      const int32_t pf = __shfl_up(f, o);
      if (lane >= o && !f) {
        sc = gcat<Program>(pp, sc);
        f = pf;
      }
    }
    if (lane == kWave - 1) {
      s_wend[wid] = sc;
      s_whead[wid] = f;
    }
    __syncthreads();
    if (!f && wid > 0) {  // the segment began in an earlier wave
      int w0 = wid - 1;
      while (w0 > 0 && !s_whead[w0]) w0--;
      P c = s_wend[w0];
      for (int w = w0 + 1; w < wid; w++) c = gcat<Program>(c, s_wend[w]);
      sc = gcat<Program>(c, sc);
    }
    s_incl[tid] = sc;
    __syncthreads();

    // Step 5: complete the first and last runs of this thread.
    if (has) {
      P carry_in{G(), 0};  // partial of item fs from the threads before this one
      if (e0 > 0 && s_slot[e0 - 1] == fs) carry_in = s_incl[tid - 1];
      const bool last_ends = e1 == ne || s_slot[e1] != ls;
      auto emit = [&](int32_t s, const P &tot) {
        const bool cont = s_offs[s + 1] > j1;
        if (s == 0) {  // began before the tile
          if (cont) {
            s_carry.open = tot.val;
            s_carry.open_has = tot.has;
            s_carry.open_i = i0 - 1;
          } else {
            s_carry.tail = tot.val;
            s_carry.tail_has = tot.has;
            s_carry.tail_i = i0 - 1;
          }
        } else if (cont) {
          s_carry.open = tot.val;
          s_carry.open_has = tot.has;
          s_carry.open_i = i0 - 1 + s;
        } else {
          op.template done<ACCUM>(s_vert[s], tot.has, tot.val);
        }
      };
      if (fs != ls) {
        emit(fs, gcat<Program>(carry_in, first));
        if (last_ends) emit(ls, run);
      } else if (last_ends) {
        emit(ls, gcat<Program>(carry_in, run));
      }
    }
    // Items with no edges in the tile: items without gather edges, and an
    // item whose edges start exactly at the tile's end (open, empty).
    for (int k = 1 + tid; k <= ni; k += kExpandBlock) {
      if (s_offs[k + 1] == s_offs[k]) {
        if (i0 - 1 + k < nf) op.template done<ACCUM>(s_vert[k], false, G());
      } else if (s_offs[k] == j1) {
        s_carry.open_has = 0;
        s_carry.open_i = i0 - 1 + k;
      }
    }
    __syncthreads();
    if (tid == 0) carry[t] = s_carry;
    __syncthreads();
  }
}

// Items that span tiles: the tile holding an item's end folds the open
// partials of the tiles before it (from the item's first tile on) and its
// own tail.
template <class Program, bool ACCUM>
__global__ void __launch_bounds__(256)
k_lbs_gather_fixup(GatherRows<Program> op, const int32_t *__restrict__ frontier, const LbsInfo *__restrict__ info,
                   const GatherCarry<typename Program::GatherType> *__restrict__ carry) {
  using G = typename Program::GatherType;
  const int64_t ntiles = (info->path + kGatherTile - 1) / kGatherTile;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < ntiles; t += (int64_t)gridDim.x * blockDim.x) {
    const int32_t it = carry[t].tail_i;
    if (it < 0) continue;
    int64_t ts = t - 1;  // carry[t - 1].open_i == it: the item was open at the previous tile's end
    while (ts > 0 && carry[ts - 1].open_i == it) ts--;
    GPart<G> acc{G(), 0};
    for (int64_t q = ts; q < t; q++) acc = gcat<Program>(acc, GPart<G>{carry[q].open, carry[q].open_has});
    acc = gcat<Program>(acc, GPart<G>{carry[t].tail, carry[t].tail_has});
    op.template done<ACCUM>(frontier[it], acc.has, acc.val);
  }
}

// Tiles a gather over at most max_path merge-path positions can need.
inline int64_t gather_carry_entries(int64_t max_path) { return max_path / kGatherTile + 2; }

}  // namespace GASengine
