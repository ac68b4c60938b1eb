// Frontier expansion kernels for gfx950, generic over an edge operation.
//
// Two strategies, as in the reference (enactor_vertex_centric.cuh:3637-3667
// picks by frontier size against `threshold`):
//
//  * Load-balance search (LBS), the role of the reference's two-phase
//    "mgpu" expand (mgpukernel/kernel.cuh:63-180 + moderngpu Scan /
//    MergePathPartitions / CTALoadBalance):
//      k_lbs_count      per frontier item: prologue + degree, workgroup-local
//                       exclusive scan, one partial per 1024 items
//      k_lbs_scan       one workgroup scans the partials -> global offsets
//      k_lbs_partition  the merge path of (items + edges) is cut into tiles
//                       of 512 / 1024 / 2048 positions (run-time choice);
//                       one thread per tile boundary
//                       finds its (item, edge) split by binary search
//      k_lbs_expand     each workgroup stages its tile's items in LDS, builds
//                       an edge->item map with a per-thread serial merge,
//                       then walks its edges with consecutive lanes on
//                       consecutive edges (coalesced column reads), whatever
//                       the degree skew.
//  * Tiered dynamic expand, the role of expand_atomic (expand_atomic/cta.cuh:
//    264-417, 701-754), re-tiled for wave64: degree >= block_degree -> whole
//    workgroup per vertex, >= wave_degree -> one wave per vertex, rest -> an
//    LDS-staged workgroup gather.  One launch, no scan: for small frontiers.
//
// Output: every edge for which op.edge() returns true appends its value to
// the next queue with one atomic per wave (wave_append).
//
// Op concept (device, passed by value):
//   void    prologue(int32_t v)          once per frontier item before degree
//   int32_t degree(int32_t v)            edges to expand (0 = none)
//   int32_t row_begin(int32_t v)         first edge slot of v
//   int32_t item_value(int32_t v)        per-item value staged in LDS and
//                                        handed to edge() (e.g. dist[v])
//   bool    edge(int32_t v, int32_t val, int32_t slot, int32_t &out)
//   void    finish()                     once per thread at kernel end (all
//                                        threads; per-thread accumulators)
// Optional staged form (k_lbs_expand evaluates a thread's VT edges of a tile
// in two passes, so the loads of all VT edges are in flight together; tail
// slots are clamped to the tile's last edge and their loads discarded):
//   using Loads = ...                    what one edge's loads produce
//   Loads load(int32_t v, int32_t val, int32_t slot)   loads only, no effects
//   bool  apply(int32_t v, Loads l, int32_t &out)      as edge(), given l
//   void  emit(bool keep, int32_t out, int32_t *s_out, int32_t *s_nout)
//         optional: what becomes of apply()'s result (whole-wave call;
//         default: wave_append to the tile's output list)
#pragma once

#include <type_traits>
#include <utility>

#include "util.h"

namespace GASengine {

template <class Op, class = void>
struct StagedEdge {
  static constexpr bool value = false;
};
template <class Op>
struct StagedEdge<Op, decltype((void)sizeof(typename Op::Loads))> {
  static constexpr bool value = true;
};
// An op may route apply()'s results itself: void emit(bool keep, int32_t
// out, int32_t *s_out, int32_t *s_nout), called by every lane of the wave
// (the default appends kept values to the tile's output list).
template <class Op, class = void>
struct EdgeEmit : std::false_type {};
template <class Op>
struct EdgeEmit<Op, std::void_t<decltype(std::declval<Op &>().emit(false, 0, (int32_t *)nullptr, (int32_t *)nullptr))>>
    : std::true_type {};
template <class Op, bool = StagedEdge<Op>::value>
struct StagedEdgeLoads {
  using type = int32_t;
};
template <class Op>
struct StagedEdgeLoads<Op, true> {
  using type = typename Op::Loads;
};

// One item per thread, so a wave's loads and stores of per-vertex arrays (an
// op's prologue touches several) cover 64 consecutive items, not 64 items
// four apart.
constexpr int kCountBlock = 1024;
constexpr int kCountIpt = 1;
constexpr int kCountTile = kCountBlock * kCountIpt;  // items per scan partial
// k_lbs_count strides over its tiles; captured loops launch it at its largest
// size every iteration, and a gated launch costs by the waves it starts:
// 512 workgroups of 1024 threads fill the GPU's wave slots once.
constexpr int kCountGridMax = 512;
constexpr int kScanBlock = 1024;
constexpr int kExpandBlock = 256;
constexpr int kExpandVt = 4;  // default 1024-position tiles: ~21 KB LDS, 7 workgroups per CU
constexpr int kLbsTile = kExpandBlock * kExpandVt;  // default merge-path positions per tile
// Compiled tile shapes (positions per workgroup tile = kExpandBlock * VT);
// lbs_launch picks one at run time (the tuner's `lbs_tile`).
constexpr int kExpandVtMin = 2, kExpandVtMax = 8;
constexpr int kLbsTileMin = kExpandBlock * kExpandVtMin;
inline bool lbs_tile_supported(int tile) { return tile == 512 || tile == 1024 || tile == 2048; }

// Device-side counters of one expansion (written by the kernels).
struct LbsInfo {
  int32_t edges;  // total edges of the frontier (sum of degrees)
  int32_t pad;
  int64_t path;   // merge path length = items + edges
};

// Partition-array entries needed for a frontier of at most n items over m
// edges, for every compiled tile shape.
inline size_t lbs_partition_entries(int64_t n, int64_t m) { return (size_t)((n + m) / kLbsTileMin + 2); }

template <class Op, bool PROLOGUE>
__global__ void __launch_bounds__(kCountBlock)
k_lbs_count(Op op, const int32_t *__restrict__ frontier, const int32_t *__restrict__ fsize,
            int32_t *__restrict__ scan_local, int32_t *__restrict__ partials) {
  __shared__ int32_t smem[kCountBlock / kWave + 1];
  const int32_t nf = *fsize;
  for (int64_t b = blockIdx.x; b * kCountTile < nf; b += gridDim.x) {
    const int64_t first = b * kCountTile + (int64_t)threadIdx.x * kCountIpt;
    int32_t deg[kCountIpt];
    int32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kCountIpt; k++) {
      deg[k] = 0;
      if (first + k < nf) {
        int32_t v = frontier[first + k];
        if (PROLOGUE) op.prologue(v);
        deg[k] = op.degree(v);
      }
      sum += deg[k];
    }
    int32_t total;
    int32_t ex = block_exclusive_scan<kCountBlock>(sum, smem, &total);
#pragma unroll
    for (int k = 0; k < kCountIpt; k++) {
      if (first + k < nf) scan_local[first + k] = ex;
      ex += deg[k];
    }
    if (threadIdx.x == 0) partials[b] = total;
  }
}

// One workgroup: exclusive scan of the per-1024-item partials in place, then
// the totals into *info and the running edge counter.
template <int BLOCK = kScanBlock>
__global__ void __launch_bounds__(BLOCK)
k_lbs_scan(const int32_t *__restrict__ fsize, int32_t *__restrict__ partials, LbsInfo *__restrict__ info,
           unsigned long long *__restrict__ edge_counter) {
  __shared__ int32_t smem[BLOCK / kWave + 1];
  const int32_t nf = *fsize;
  const int32_t nb = (int32_t)(((int64_t)nf + kCountTile - 1) / kCountTile);
  int32_t carry = 0;
  for (int32_t base = 0; base < nb; base += BLOCK) {
    int32_t i = base + threadIdx.x;
    int32_t v = i < nb ? partials[i] : 0;
    int32_t tot;
    int32_t ex = block_exclusive_scan<BLOCK>(v, smem, &tot);
    if (i < nb) partials[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) {
    info->edges = carry;
    info->path = (int64_t)nf + carry;
    if (edge_counter) atomicAdd(edge_counter, (unsigned long long)carry);
  }
}

__device__ __forceinline__ int32_t lbs_offset(const int32_t *scan_local, const int32_t *partials, int32_t i) {
  return scan_local[i] + partials[i / kCountTile];
}

// Items consumed at merge-path diagonal d (items precede edges on ties).
__device__ __forceinline__ int32_t lbs_path_search(const int32_t *scan_local, const int32_t *partials, int64_t d,
                                                   int32_t nf, int32_t ne) {
  int64_t lo = d - ne > 0 ? d - ne : 0;
  int64_t hi = d < nf ? d : nf;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if ((int64_t)lbs_offset(scan_local, partials, (int32_t)mid) <= d - mid - 1)
      lo = mid + 1;
    else
      hi = mid;
  }
  return (int32_t)lo;
}

// parts[t] = items consumed at diagonal min(t * TILE, path), t = 0..ntiles.
template <int TILE = kLbsTile>
__global__ void __launch_bounds__(256)
k_lbs_partition(const int32_t *__restrict__ fsize, const int32_t *__restrict__ scan_local,
                const int32_t *__restrict__ partials, const LbsInfo *__restrict__ info, int32_t *__restrict__ parts) {
  const int32_t nf = *fsize;
  const int32_t ne = info->edges;
  const int64_t path = info->path;
  const int64_t ntiles = (path + TILE - 1) / TILE;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t <= ntiles; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t d = t * TILE < path ? t * TILE : path;
    parts[t] = lbs_path_search(scan_local, partials, d, nf, ne);
  }
}

template <class Op, int VT = kExpandVt>
__global__ void __launch_bounds__(kExpandBlock)
k_lbs_expand(Op op, const int32_t *__restrict__ frontier, const int32_t *__restrict__ fsize,
             const int32_t *__restrict__ scan_local, const int32_t *__restrict__ partials,
             const LbsInfo *__restrict__ info, const int32_t *__restrict__ parts, int32_t *__restrict__ out_queue,
             int32_t *__restrict__ out_size) {
  constexpr int TILE = kExpandBlock * VT;
  // slot k <-> frontier item i0 - 1 + k (slot 0: the item owning the tile's first edges)
  __shared__ int32_t s_offs[TILE + 1];
  __shared__ int32_t s_vert[TILE + 1];
  __shared__ int32_t s_begin[TILE + 1];
  __shared__ int32_t s_val[Op::kItemValues ? TILE + 1 : 1];
  __shared__ int32_t s_slot[TILE];
  __shared__ int32_t s_out[TILE];  // this tile's outputs, flushed with one global atomic
  __shared__ int32_t s_nout, s_obase;
  const int64_t path = info->path;
  for (int64_t t = blockIdx.x; t * TILE < path; t += gridDim.x) {
    const int64_t d0 = t * TILE;
    const int64_t d1 = d0 + TILE < path ? d0 + TILE : path;
    const int32_t i0 = parts[t], i1 = parts[t + 1];
    const int32_t j0 = (int32_t)(d0 - i0), j1 = (int32_t)(d1 - i1);
    const int32_t ni = i1 - i0, ne = j1 - j0;
    for (int k = threadIdx.x; k <= ni; k += kExpandBlock) {
      const int32_t i = i0 - 1 + k;
      if (i >= 0) {
        const int32_t v = frontier[i];
        s_vert[k] = v;
        s_offs[k] = lbs_offset(scan_local, partials, i);
        s_begin[k] = op.row_begin(v);
        if (Op::kItemValues) s_val[k] = op.item_value(v);
      } else {
        s_vert[k] = -1;
        s_offs[k] = 0;
        s_begin[k] = 0;
        if (Op::kItemValues) s_val[k] = 0;
      }
    }
    if (threadIdx.x == 0) s_nout = 0;
    __syncthreads();
    {
      const int32_t dl = (int32_t)threadIdx.x * VT;
      if (dl < ni + ne) {
        int32_t lo = dl - ne > 0 ? dl - ne : 0, hi = dl < ni ? dl : ni;
        while (lo < hi) {
          int32_t mid = (lo + hi) >> 1;
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
    if constexpr (StagedEdge<Op>::value) {
      if (ne > 0) {  // ne <= TILE = VT * kExpandBlock: one pass
        typename StagedEdgeLoads<Op>::type pre[VT];
        int32_t sl[VT];
#pragma unroll
        for (int r = 0; r < VT; r++) {
          const int32_t k = min((int32_t)threadIdx.x + r * kExpandBlock, ne - 1);
          const int32_t s = s_slot[k];
          sl[r] = s;
          pre[r] = op.load(s_vert[s], Op::kItemValues ? s_val[s] : 0, s_begin[s] + (j0 + k - s_offs[s]));
        }
#pragma unroll
        for (int r = 0; r < VT; r++) {
          const int32_t k = (int32_t)threadIdx.x + r * kExpandBlock;
          bool keep = false;
          int32_t out = 0;
          if (k < ne) keep = op.apply(s_vert[sl[r]], pre[r], out);
          if constexpr (EdgeEmit<Op>::value)
            op.emit(keep, out, s_out, &s_nout);
          else
            wave_append(keep, out, s_out, &s_nout);
        }
      }
    } else {
      for (int32_t base = 0; base < ne; base += kExpandBlock) {
        const int32_t k = base + threadIdx.x;
        bool keep = false;
        int32_t out = 0;
        if (k < ne) {
          const int32_t s = s_slot[k];
          keep = op.edge(s_vert[s], Op::kItemValues ? s_val[s] : 0, s_begin[s] + (j0 + k - s_offs[s]), out);
        }
        wave_append(keep, out, s_out, &s_nout);  // LDS atomics, one per wave
      }
    }
    __syncthreads();
    const int32_t nout = s_nout;
    if (nout) {
      if (threadIdx.x == 0) s_obase = atomicAdd(out_size, nout);
      __syncthreads();
      const int32_t ob = s_obase;
      for (int32_t k = threadIdx.x; k < nout; k += kExpandBlock) out_queue[ob + k] = s_out[k];
      __syncthreads();
    }
  }
  op.finish();
}

template <class Op, bool PROLOGUE>
__global__ void __launch_bounds__(kExpandBlock)
k_tiered_expand(Op op, const int32_t *__restrict__ frontier, const int32_t *__restrict__ fsize,
                int32_t *__restrict__ out_queue, int32_t *__restrict__ out_size, int32_t wave_degree,
                int32_t block_degree, unsigned long long *__restrict__ edge_counter) {
  constexpr int B = kExpandBlock;
  __shared__ int32_t s_offs[B];
  __shared__ int32_t s_vert[B];
  __shared__ int32_t s_begin[B];
  __shared__ int32_t s_val[B];
  __shared__ int32_t s_scan[B / kWave + 1];
  __shared__ int32_t s_owner;
  __shared__ int32_t s_big[4];
  __shared__ unsigned long long s_red[B / kWave];
  const int32_t nf = *fsize;
  unsigned long long inspected = 0;
  for (int64_t base = (int64_t)blockIdx.x * B; base < nf; base += (int64_t)gridDim.x * B) {
    const int64_t i = base + threadIdx.x;
    int32_t v = -1, deg = 0, begin = 0, val = 0;
    if (i < nf) {
      v = frontier[i];
      if (PROLOGUE) op.prologue(v);
      deg = op.degree(v);
      if (deg > 0) {
        begin = op.row_begin(v);
        val = op.item_value(v);
      }
    }
    inspected += (unsigned long long)deg;
    // Tier 1: whole workgroup per high-degree vertex.
    for (;;) {
      if (threadIdx.x == 0) s_owner = -1;
      __syncthreads();
      if (deg >= block_degree) s_owner = threadIdx.x;
      __syncthreads();
      const int owner = s_owner;
      if (owner < 0) break;
      if ((int)threadIdx.x == owner) {
        s_big[0] = v;
        s_big[1] = begin;
        s_big[2] = deg;
        s_big[3] = val;
        deg = 0;
      }
      __syncthreads();
      const int32_t bv = s_big[0], bb = s_big[1], bd = s_big[2], bval = s_big[3];
      for (int32_t k0 = 0; k0 < bd; k0 += B) {
        const int32_t k = k0 + threadIdx.x;
        bool keep = false;
        int32_t out = 0;
        if (k < bd) keep = op.edge(bv, bval, bb + k, out);
        wave_append(keep, out, out_queue, out_size);
      }
      __syncthreads();
    }
    // Tier 2: one wave per mid-degree vertex.
    uint64_t mask = ballot(deg >= wave_degree);
    while (mask) {
      const int l = __builtin_ctzll(mask);
      mask &= mask - 1;
      const int32_t wv = __shfl(v, l), wb = __shfl(begin, l), wd = __shfl(deg, l), wval = __shfl(val, l);
      for (int32_t k0 = 0; k0 < wd; k0 += kWave) {
        const int32_t k = k0 + lane_id();
        bool keep = false;
        int32_t out = 0;
        if (k < wd) keep = op.edge(wv, wval, wb + k, out);
        wave_append(keep, out, out_queue, out_size);
      }
    }
    if (deg >= wave_degree) deg = 0;
    // Tier 3: LDS-staged gather of the remaining low-degree edges.
    int32_t total;
    const int32_t ex = block_exclusive_scan<B>(deg, s_scan, &total);
    s_offs[threadIdx.x] = ex;
    s_vert[threadIdx.x] = v;
    s_begin[threadIdx.x] = begin;
    s_val[threadIdx.x] = val;
    __syncthreads();
    for (int32_t k0 = 0; k0 < total; k0 += B) {
      const int32_t k = k0 + threadIdx.x;
      bool keep = false;
      int32_t out = 0;
      if (k < total) {
        int32_t lo = 0, hi = B;
        while (lo < hi) {
          int32_t mid = (lo + hi) >> 1;
          if (s_offs[mid] <= k)
            lo = mid + 1;
          else
            hi = mid;
        }
        const int32_t j = lo - 1;
        keep = op.edge(s_vert[j], s_val[j], s_begin[j] + (k - s_offs[j]), out);
      }
      wave_append(keep, out, out_queue, out_size);
    }
    __syncthreads();
  }
  if (edge_counter) {
    unsigned long long t = block_sum_u64<B>(inspected, s_red);
    if (threadIdx.x == 0 && t) atomicAdd(edge_counter, t);
  }
  op.finish();
}

// Per-frontier-item prologue only (apply/post_apply) for the tiered path,
// which must finish before any contract of the same iteration runs.
template <class Op>
__global__ void __launch_bounds__(256)
k_frontier_prologue(Op op, const int32_t *__restrict__ frontier, const int32_t *__restrict__ fsize) {
  const int32_t nf = *fsize;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nf; i += (int64_t)gridDim.x * blockDim.x)
    op.prologue(frontier[i]);
}

// ---- pull over all vertices: merge path of (n vertices, m in-edges) -------
//
// Items are the vertices 0..n-1 themselves, so their edge offsets are the
// CSC column_offsets array and the tile partition depends only on the graph:
// k_static_partition runs once per graph.  k_pull_reduce maps each tile's
// edges to vertex slots like k_lbs_expand, folds per-edge values into an LDS
// accumulator per vertex, and reports each vertex's result with a flag
// telling whether all of its in-edges fell inside the tile (then the tile is
// the only writer of that vertex and may store plainly).
//
// PullOp concept:
//   T    identity()
//   T    edge(int32_t v, int32_t slot)             value of in-edge slot of v
//   void combine(T *lds_acc, T val)               LDS reduction (e.g. atomicMin)
//   void item_done(int32_t v, T acc, bool whole)  result of v for this tile
//   void finish()
template <int DUMMY = 0>
__global__ void __launch_bounds__(256)
k_static_partition(const int32_t *__restrict__ offs, int32_t n, int64_t m, int32_t *__restrict__ parts) {
  const int64_t path = (int64_t)n + m;
  const int64_t ntiles = (path + kLbsTile - 1) / kLbsTile;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t <= ntiles; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t d = t * kLbsTile < path ? t * kLbsTile : path;
    int64_t lo = d - m > 0 ? d - m : 0, hi = d < n ? d : n;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if ((int64_t)offs[mid] <= d - mid - 1)
        lo = mid + 1;
      else
        hi = mid;
    }
    parts[t] = (int32_t)lo;
  }
}

// skip: optional device flag; the kernel returns at once when it is nonzero
// (lets a captured launch sequence switch the pull off).
template <class Op, class T>
__global__ void __launch_bounds__(kExpandBlock)
k_pull_reduce(Op op, const int32_t *__restrict__ offs, int32_t n, int64_t m, const int32_t *__restrict__ parts,
              const int32_t *__restrict__ skip) {
  if (skip && *skip) return;
  constexpr int TILE = kLbsTile;
  constexpr int VT = kExpandVt;
  __shared__ int32_t s_offs[TILE + 2];  // s_offs[k] = offs[i0 - 1 + k], k = 0..ni+1
  __shared__ T s_acc[TILE + 1];
  __shared__ int32_t s_slot[TILE];
  const int64_t path = (int64_t)n + m;
  for (int64_t t = blockIdx.x; t * TILE < path; t += gridDim.x) {
    const int64_t d0 = t * TILE;
    const int64_t d1 = d0 + TILE < path ? d0 + TILE : path;
    const int32_t i0 = parts[t], i1 = parts[t + 1];
    const int32_t j0 = (int32_t)(d0 - i0), j1 = (int32_t)(d1 - i1);
    const int32_t ni = i1 - i0, ne = j1 - j0;
    for (int k = threadIdx.x; k <= ni + 1; k += kExpandBlock) {
      const int32_t i = i0 - 1 + k;
      s_offs[k] = i < 0 ? 0 : offs[i];  // i <= i1 <= n: offs has n + 1 entries
      if (k <= ni) s_acc[k] = op.identity();
    }
    __syncthreads();
    {
      const int32_t dl = (int32_t)threadIdx.x * VT;
      if (dl < ni + ne) {
        int32_t lo = dl - ne > 0 ? dl - ne : 0, hi = dl < ni ? dl : ni;
        while (lo < hi) {
          int32_t mid = (lo + hi) >> 1;
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
    for (int32_t k = threadIdx.x; k < ne; k += kExpandBlock) {
      const int32_t s = s_slot[k];
      const T val = op.edge(i0 - 1 + s, j0 + k);
      op.combine(&s_acc[s], val);
    }
    __syncthreads();
    for (int k = threadIdx.x; k <= ni; k += kExpandBlock) {
      const int32_t v = i0 - 1 + k;
      if (v < 0) continue;
      const bool whole = k >= 1 && s_offs[k] >= j0 && s_offs[k + 1] <= j1;
      op.item_done(v, s_acc[k], whole);
    }
    __syncthreads();
  }
  op.finish();
}

// ---- staged pull: two dependent memory round trips per tile -------------
//
// k_pull_reduce waits on the tile's offsets, then its edges' loads, then
// their gathers, then each vertex's current state.  Here every load that
// does not depend on a gather is issued at once (stage 1; the tile's edges
// are the contiguous slots [j0, j1), known from the partition alone): the
// row offsets, each thread's VT edges (Edge loads) and the current state of
// the tile's vertices.  When a tile starts, its VT gathers go out together
// (stage 2: they need only the edges, already in registers), the merge-path
// slot map is built in LDS while they are in flight, and the results are
// combined and handed to item_done with the prefetched state.
//
// The loop is software-pipelined: a workgroup issues the stage-1 loads of
// its next tile (and the partition bounds of the one after) before the
// item_done writes of the current tile.  gfx950 counts loads and stores in
// one in-order counter, so loads issued after a tile's stores would wait
// for those stores as well.  The next tile (t + gridDim.x) shares no vertex
// that the current tile writes with a plain store: such a vertex has all of
// its in-edges in the current tile.
//
// StagedPullOp concept:
//   using Edge = ...                        per-edge loads (e.g. source and weight)
//   Edge    load_edge(int32_t slot)         loads only
//   T       gather(const Edge &e)           dependent loads (e.g. dist[src] + w)
//   void    combine(T *lds_acc, T val)
//   int32_t vertex_state(int32_t v)         prefetched per-vertex value
//   T       acc_start(int32_t state)        a row's accumulator before its
//                                           gathers (for a minimum: the
//                                           vertex's own value, so gathers
//                                           that cannot improve it skip LDS)
//   void    item_done(int32_t v, T acc, bool whole, int32_t state)
//           whole: every in-edge of v lies in this tile, so no other
//           workgroup writes v during the launch and `state` is exact
//   void    finish()
// Optional row filter (an op that knows which rows cannot change):
//   int32_t row_limit()                     read once per launch; a row whose
//                                           prefetched state is <= it has no
//                                           gathers (its edges still stream)
//   T       skipped()                       the value combined for them
// With a filter the gathers go out after the slot map (they need their
// row's state) instead of before it.
// Optional tile filter:
//   const int32_t *row_prefix()             prefix count over the vertices of
//                                           the rows that can change (n + 1
//                                           entries); a tile whose rows
//                                           [i0 - 1, i1) hold none is skipped
//                                           whole, its edges never loaded
//   void    tile_visited(int32_t edges, int32_t items)   per tile visited,
//                                           called by one thread
template <class Op, class = void>
struct PullRowFilter : std::false_type {};
template <class Op>
struct PullRowFilter<Op, std::void_t<decltype(std::declval<const Op &>().row_limit())>> : std::true_type {};
template <class Op, class = void>
struct PullTileFilter : std::false_type {};
template <class Op>
struct PullTileFilter<Op, std::void_t<decltype(std::declval<const Op &>().row_prefix())>> : std::true_type {};

template <class Op, class T>
__global__ void __launch_bounds__(kExpandBlock)
k_pull_staged(Op op, const int32_t *__restrict__ offs, int32_t n, int64_t m, const int32_t *__restrict__ parts,
              const int32_t *__restrict__ skip) {
  if (skip && *skip) return;
  constexpr int TILE = kLbsTile;
  constexpr int VT = kExpandVt;
  constexpr int Q = (TILE + 2 + kExpandBlock - 1) / kExpandBlock;  // per-thread share of the offsets
  __shared__ int32_t s_offs[TILE + 2];  // s_offs[k] = offs[i0 - 1 + k], k = 0..ni+1
  __shared__ T s_acc[TILE + 1];
  __shared__ int32_t s_state[TILE + 1];
  __shared__ int32_t s_slot[TILE];
  const int tid = threadIdx.x;
  const int64_t path = (int64_t)n + m;
  const int64_t G = gridDim.x;
  typename Op::Edge ed[VT];
  int32_t ov[Q], sv[Q];
  int32_t limit = 0;
  if constexpr (PullRowFilter<Op>::value) limit = op.row_limit();
  const int32_t *rows = nullptr;
  if constexpr (PullTileFilter<Op>::value) rows = op.row_prefix();
  // Stage 1 of tile tt with partition bounds p0, p1 into ed / ov / sv; false
  // (nothing loaded) when the tile filter finds no row that can change.
  auto stage1 = [&](int64_t tt, int32_t p0, int32_t p1) -> bool {
    if constexpr (PullTileFilter<Op>::value) {
      if (rows[p1] == rows[p0 > 0 ? p0 - 1 : 0]) return false;
    }
    const int64_t e0 = tt * TILE;
    const int64_t e1 = e0 + TILE < path ? e0 + TILE : path;
    const int32_t jj0 = (int32_t)(e0 - p0), nne = (int32_t)(e1 - p1) - jj0, nni = p1 - p0;
    if (nne > 0) {
#pragma unroll
      for (int r = 0; r < VT; r++) ed[r] = op.load_edge(jj0 + min(tid + r * kExpandBlock, nne - 1));
    }
#pragma unroll
    for (int q = 0; q < Q; q++) {
      const int32_t k = min(tid + q * kExpandBlock, nni + 1);
      const int32_t i = p0 - 1 + k;  // <= p1 <= n: offs has n + 1 entries
      ov[q] = i < 0 ? 0 : offs[i];
      sv[q] = (i < 0 || i >= n) ? 0 : op.vertex_state(i);
    }
    return true;
  };
  int64_t t = blockIdx.x;
  int32_t a0 = 0, a1 = 0, b0 = 0, b1 = 0;  // partition bounds of tiles t and t + G
  bool act = false;                        // tile t's stage 1 ran (the tile filter kept it)
  if (t * TILE < path) {
    a0 = parts[t];
    a1 = parts[t + 1];
    act = stage1(t, a0, a1);
  }
  if ((t + G) * TILE < path) {
    b0 = parts[t + G];
    b1 = parts[t + G + 1];
  }
  for (; t * TILE < path; t += G) {
    if (!act) {  // (uniform) a tile of settled rows only: on to the next one
      bool act_next = false;
      int32_t c0 = 0, c1 = 0;
      if ((t + G) * TILE < path) {
        act_next = stage1(t + G, b0, b1);
        if ((t + 2 * G) * TILE < path) {
          c0 = parts[t + 2 * G];
          c1 = parts[t + 2 * G + 1];
        }
      }
      a0 = b0;
      a1 = b1;
      b0 = c0;
      b1 = c1;
      act = act_next;
      continue;
    }
    const int64_t d0 = t * TILE;
    const int64_t d1 = d0 + TILE < path ? d0 + TILE : path;
    const int32_t i0 = a0, i1 = a1;
    const int32_t j0 = (int32_t)(d0 - i0), j1 = (int32_t)(d1 - i1);
    const int32_t ni = i1 - i0, ne = j1 - j0;
    // Stage 2 first: the gathers depend only on the edges already in
    // registers, so they are in flight while the slot map is built (without
    // a row filter).
    T val[VT];
    if constexpr (!PullRowFilter<Op>::value) {
      if (ne > 0) {
#pragma unroll
        for (int r = 0; r < VT; r++) val[r] = op.gather(ed[r]);
      }
    }
#pragma unroll
    for (int q = 0; q < Q; q++) {
      const int32_t k = tid + q * kExpandBlock;
      if (k <= ni + 1) s_offs[k] = ov[q];
      if (k <= ni) {
        s_state[k] = sv[q];
        s_acc[k] = op.acc_start(sv[q]);
      }
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
    if constexpr (PullRowFilter<Op>::value) {
      // Gathers of the rows that can still change only.
      if (ne > 0) {
#pragma unroll
        for (int r = 0; r < VT; r++) {
          const int32_t k = tid + r * kExpandBlock;
          val[r] = k < ne && s_state[s_slot[k]] > limit ? op.gather(ed[r]) : op.skipped();
        }
      }
    }
    // Combine the gathered values.
    if (ne > 0) {
#pragma unroll
      for (int r = 0; r < VT; r++) {
        const int32_t k = tid + r * kExpandBlock;
        if (k < ne) op.combine(&s_acc[s_slot[k]], val[r]);
      }
    }
    __syncthreads();
    if constexpr (PullTileFilter<Op>::value) {
      if (tid == 0) op.tile_visited(ne, ni + 1);
    }
    // Next tile's stage 1 and the bounds of the one after, ahead of this
    // tile's writes.
    int32_t c0 = 0, c1 = 0;
    bool act_next = false;
    if ((t + G) * TILE < path) {
      act_next = stage1(t + G, b0, b1);
      if ((t + 2 * G) * TILE < path) {
        c0 = parts[t + 2 * G];
        c1 = parts[t + 2 * G + 1];
      }
    }
    for (int k = tid; k <= ni; k += kExpandBlock) {
      const int32_t v = i0 - 1 + k;
      if (v < 0) continue;
      const bool whole = k >= 1 && s_offs[k] >= j0 && s_offs[k + 1] <= j1;
      op.item_done(v, s_acc[k], whole, s_state[k]);
    }
    __syncthreads();
    a0 = b0;
    a1 = b1;
    b0 = c0;
    b1 = c1;
    act = act_next;
  }
  op.finish();
}

// Launch the four LBS kernels on stream s: count (with or without the
// prologue), scan, partition, expand.  `grid_items` = host-known frontier
// size (for grid sizing; kernels read the device-side size).  `max_path`
// bounds items + edges for the partition grid.
struct LbsBuffers {
  int32_t *scan_local;
  int32_t *partials;
  LbsInfo *info;
  int32_t *parts;
  unsigned long long *edge_counter;
  int expand_grid;        // persistent workgroups of k_lbs_expand
  int vt = kExpandVt;     // positions per thread: tile = kExpandBlock * vt (2, 4 or 8)
};

template <int VT, class Op, class Timer>
inline void lbs_partition_expand(const Op &op, const int32_t *frontier, const int32_t *fsize, int64_t max_path,
                                 const LbsBuffers &b, int32_t *out_queue, int32_t *out_size, hipStream_t s,
                                 Timer &timer) {
  constexpr int TILE = kExpandBlock * VT;
  hipLaunchKernelGGL(k_lbs_partition<TILE>, dim3(grid_for(max_path / TILE + 2, 256, 4096)), dim3(256), 0, s, fsize,
                     b.scan_local, b.partials, b.info, b.parts);
  timer(3);
  hipLaunchKernelGGL((k_lbs_expand<Op, VT>), dim3(b.expand_grid), dim3(kExpandBlock), 0, s, op, frontier, fsize,
                     b.scan_local, b.partials, b.info, b.parts, out_queue, out_size);
}

template <class Op, class Timer>
inline void lbs_launch(const Op &op, bool prologue, const int32_t *frontier, const int32_t *fsize, int64_t nf_host,
                       int64_t max_path, const LbsBuffers &b, int32_t *out_queue, int32_t *out_size, hipStream_t s,
                       Timer &&timer) {
  timer(0);
  if (prologue)
    hipLaunchKernelGGL((k_lbs_count<Op, true>), dim3(grid_for(nf_host, kCountTile, kCountGridMax)), dim3(kCountBlock), 0, s, op,
                       frontier, fsize, b.scan_local, b.partials);
  else
    hipLaunchKernelGGL((k_lbs_count<Op, false>), dim3(grid_for(nf_host, kCountTile, kCountGridMax)), dim3(kCountBlock), 0, s, op,
                       frontier, fsize, b.scan_local, b.partials);
  timer(1);
  hipLaunchKernelGGL(k_lbs_scan<kScanBlock>, dim3(1), dim3(kScanBlock), 0, s, fsize, b.partials, b.info,
                     b.edge_counter);
  timer(2);
  if (b.vt == 2)
    lbs_partition_expand<2>(op, frontier, fsize, max_path, b, out_queue, out_size, s, timer);
  else if (b.vt == 8)
    lbs_partition_expand<8>(op, frontier, fsize, max_path, b, out_queue, out_size, s, timer);
  else
    lbs_partition_expand<4>(op, frontier, fsize, max_path, b, out_queue, out_size, s, timer);
  timer(4);
}

}  // namespace GASengine
