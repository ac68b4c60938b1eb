// Load-balanced, deterministic row sums: a merge-path segmented reduction
// (PageRank sweeps over CSC in-edges, its adjoint over CSR out-edges).
//
// The merge path of (n row starts, m edges) is cut into tiles of kSegTile
// positions by the static partition (k_static_partition, the one the pull
// kernels use), so every workgroup handles the same amount of work whatever
// the degree skew: a hub row is spread over many tiles instead of being one
// workgroup's serial job.  Per tile:
//   1. the tile's row offsets are staged in LDS and each thread finds, by a
//      merge, the row of each of its VT path positions (s_slot, as in
//      k_pull_reduce);
//   2. the edge contributions are gathered with consecutive lanes on
//      consecutive edges (VT loads in flight per thread, issued when the tile
//      starts, before step 1's LDS work) into LDS;
//   3. thread t sums edges [VT t, VT t + VT) of the tile in order, one
//      Neumaier-compensated FP32 partial per run of equal rows;
//   4. a segmented inclusive scan of the threads' last-run partials (wave
//      shuffles, then the four waves in order) gives each thread the partial
//      of its first row from the threads before it;
//   5. a row that starts and ends in the tile is handed to op.item_done;
//      the row still open at the tile's end and the row that started before
//      the tile and ends in it leave their in-tile partials in the tile's
//      SegCarry, and k_seg_fixup folds those, tile by tile in path order,
//      for every row that spans tiles.
// Every fold runs in a fixed order that depends only on the graph and
// kSegTile, so results are identical run to run and for any grid size.
//
// SegRowOp concept (device, by value):
//   void  begin()                         per-thread setup at kernel start
//   using Edge = ...                      per-edge loads (e.g. source, weight)
//   Edge  load_edge(int32_t k)            loads only
//   float gather(const Edge &e)           the edge's contribution (dependent loads)
//   using State = ...                     per-vertex values item_done needs
//   State vertex_state(int32_t v)         loads only; read-only data during a launch
//   void  item_done(int32_t v, float sum, const State &st)
//   void  finish()                        all threads of every workgroup, once
// item_done is called exactly once per vertex, by k_seg_rows or k_seg_fixup.
// Per tile, the first stage (row offsets, the tile's Edge loads and its
// vertices' State) is issued together, a tile ahead: the workgroup loads its
// next tile's first stage before it writes the current tile's results
// (gfx950 counts loads and stores in one in-order counter).  The second
// stage is the VT gathers per thread.
#pragma once

#include "lbs.h"
#include "rowreduce.h"

namespace GASengine {

constexpr int kSegVt = kExpandVt;
constexpr int kSegTile = kRowBlock * kSegVt;
static_assert(kSegTile == kLbsTile, "segmented rows share the static merge-path partition");

// Per tile: partial of the row open at the tile's end (open_v, -1 if none)
// and in-tile partial of the row that began before the tile and ends in it
// (tail_v, -1 if none).
struct SegCarry {
  float2 open, tail;
  int32_t open_v, tail_v;
};

inline int64_t seg_tiles(int64_t n, int64_t m) { return (n + m + kSegTile - 1) / kSegTile; }

__device__ __forceinline__ KSum ksum_of(float s, float c) {
  KSum k;
  k.s = s;
  k.c = c;
  return k;
}
// a then b, as one compensated sum.
__device__ __forceinline__ KSum ksum_cat(KSum a, const KSum &b) {
  a.merge(b.s, b.c);
  return a;
}

template <class Op>
__global__ void __launch_bounds__(kRowBlock)
k_seg_rows(Op op, const int32_t *__restrict__ offs, int32_t n, int64_t m, const int32_t *__restrict__ parts,
           SegCarry *__restrict__ carry, const int32_t *__restrict__ done) {
  constexpr int TILE = kSegTile, VT = kSegVt, W = kRowBlock / kWave;
  constexpr int Q = (TILE + 2 + kRowBlock - 1) / kRowBlock;  // per-thread share of the offsets
  using State = typename Op::State;
  __shared__ int32_t s_offs[TILE + 2];  // s_offs[k] = offs[i0 - 1 + k], k = 0..ni+1
  __shared__ __align__(16) int32_t s_slot[TILE];
  __shared__ __align__(16) float s_val[TILE];
  __shared__ State s_state[TILE + 1];   // s_state[k]: vertex i0 - 1 + k
  __shared__ float2 s_incl[kRowBlock];
  __shared__ float2 s_wend[W];
  __shared__ int32_t s_whead[W];
  __shared__ SegCarry s_carry;
  if (done && *done) return;
  op.begin();
  const int tid = threadIdx.x, lane = lane_id(), wid = tid / kWave;
  const int64_t path = (int64_t)n + m;
  const int64_t G = gridDim.x;
  typename Op::Edge ed[VT];
  int32_t ov[Q];
  State sv[Q];
  // First stage of tile tt (partition bounds p0, p1) into ed / ov / sv.
  auto stage1 = [&](int64_t tt, int32_t p0, int32_t p1) {
    const int64_t e0 = tt * TILE;
    const int64_t e1 = e0 + TILE < path ? e0 + TILE : path;
    const int32_t jj0 = (int32_t)(e0 - p0), nne = (int32_t)(e1 - p1) - jj0, nni = p1 - p0;
    if (nne > 0) {
#pragma unroll
      for (int r = 0; r < VT; r++) ed[r] = op.load_edge(jj0 + min(tid + r * kRowBlock, nne - 1));
    }
#pragma unroll
    for (int q = 0; q < Q; q++) {
      const int32_t k = min(tid + q * kRowBlock, nni + 1);
      const int32_t i = p0 - 1 + k;  // <= p1 <= n: offs has n + 1 entries
      ov[q] = i < 0 ? 0 : offs[i];
      sv[q] = op.vertex_state(i < 0 ? 0 : i < n ? i : n - 1);
    }
  };
  int64_t t = blockIdx.x;
  int32_t a0 = 0, a1 = 0, b0 = 0, b1 = 0;  // partition bounds of tiles t and t + G
  if (t * TILE < path) {
    a0 = parts[t];
    a1 = parts[t + 1];
    stage1(t, a0, a1);
  }
  if ((t + G) * TILE < path) {
    b0 = parts[t + G];
    b1 = parts[t + G + 1];
  }
  for (; t * TILE < path; t += G) {
    const int64_t d0 = t * TILE;
    const int64_t d1 = d0 + TILE < path ? d0 + TILE : path;
    const int32_t i0 = a0, i1 = a1;
    const int32_t j0 = (int32_t)(d0 - i0), j1 = (int32_t)(d1 - i1);
    const int32_t ni = i1 - i0, ne = j1 - j0;
    // Second stage first: the gathers need only the edges, already in
    // registers, so they are in flight while the slot map is built.
    float val[VT];
    if (ne > 0) {
#pragma unroll
      for (int r = 0; r < VT; r++) val[r] = op.gather(ed[r]);
    }
#pragma unroll
    for (int q = 0; q < Q; q++) {
      const int32_t k = tid + q * kRowBlock;
      if (k <= ni + 1) s_offs[k] = ov[q];
      if (k <= ni) s_state[k] = sv[q];
    }
    if (tid == 0) {
      s_carry.open_v = -1;
      s_carry.tail_v = -1;
      s_carry.open = make_float2(0.f, 0.f);
      s_carry.tail = make_float2(0.f, 0.f);
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
    if (ne > 0) {  // the gathered values (ne <= TILE)
#pragma unroll
      for (int r = 0; r < VT; r++)
        if (tid + r * kRowBlock < ne) s_val[tid + r * kRowBlock] = val[r];
    }
    __syncthreads();

    // Step 3: this thread's edges [e0, e1) in runs of equal rows.
    const int32_t e0 = tid * VT, e1 = min(e0 + VT, ne);
    const bool has = e0 < ne;
    int32_t fs = -1, ls = -1;
    KSum first, run;  // first run's partial (when the thread holds two or more runs), current run
    bool head = true;
    if (has) {
      const int4 sl4 = reinterpret_cast<const int4 *>(s_slot)[tid];
      const float4 v4 = reinterpret_cast<const float4 *>(s_val)[tid];
      const int32_t sl[VT] = {sl4.x, sl4.y, sl4.z, sl4.w};
      const float vv[VT] = {v4.x, v4.y, v4.z, v4.w};
      fs = sl[0];
      int32_t cs = fs;
      int nruns = 1;
#pragma unroll
      for (int k = 0; k < VT; k++) {
        if (e0 + k >= e1) break;
        if (sl[k] != cs) {
          if (nruns == 1) {
            first = run;
          } else {
            // A run that starts and ends inside this thread: the row's whole
            // in-tile sum.  Rows strictly inside one thread start in the tile.
            const int32_t s = cs;
            const bool cont = s_offs[s + 1] > j1;
            if (!cont && s >= 1) op.item_done(i0 - 1 + s, run.value(), s_state[s]);
          }
          run = KSum();
          cs = sl[k];
          nruns++;
        }
        run.add(vv[k]);
      }
      ls = cs;
      head = nruns > 1 || e0 == 0 || s_slot[e0 - 1] != ls;
    }

    // Step 4: segmented inclusive scan of the last-run partials.
    float ss = run.s, sc = run.c;
    int32_t f = head ? 1 : 0;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const float ps = __shfl_up(ss, o), pc = __shfl_up(sc, o);
      const int32_t pf = __shfl_up(f, o);
      if (lane >= o && !f) {
        const KSum a = ksum_cat(ksum_of(ps, pc), ksum_of(ss, sc));
        ss = a.s;
        sc = a.c;
        f = pf;
      }
    }
    if (lane == kWave - 1) {
      s_wend[wid] = make_float2(ss, sc);
      s_whead[wid] = f;
    }
    __syncthreads();
    if (!f && wid > 0) {  // the segment began in an earlier wave
      int w0 = wid - 1;
      while (w0 > 0 && !s_whead[w0]) w0--;
      KSum c = ksum_of(s_wend[w0].x, s_wend[w0].y);
      for (int w = w0 + 1; w < wid; w++) c = ksum_cat(c, ksum_of(s_wend[w].x, s_wend[w].y));
      const KSum a = ksum_cat(c, ksum_of(ss, sc));
      ss = a.s;
      sc = a.c;
    }
    s_incl[tid] = make_float2(ss, sc);
    __syncthreads();
    // Next tile's first stage (and the bounds of the one after), ahead of
    // this tile's writes.
    int32_t c0 = 0, c1 = 0;
    if ((t + G) * TILE < path) {
      stage1(t + G, b0, b1);
      if ((t + 2 * G) * TILE < path) {
        c0 = parts[t + 2 * G];
        c1 = parts[t + 2 * G + 1];
      }
    }

    // Step 5: complete the first and last runs of this thread.
    if (has) {
      KSum carry_in;  // partial of row fs from the threads before this one
      if (e0 > 0 && s_slot[e0 - 1] == fs) carry_in = ksum_of(s_incl[tid - 1].x, s_incl[tid - 1].y);
      const bool last_ends = e1 == ne || s_slot[e1] != ls;
      auto emit = [&](int32_t s, const KSum &tot) {
        const bool cont = s_offs[s + 1] > j1;
        if (s == 0) {  // began before the tile
          if (cont) {
            s_carry.open = make_float2(tot.s, tot.c);
            s_carry.open_v = i0 - 1;
          } else {
            s_carry.tail = make_float2(tot.s, tot.c);
            s_carry.tail_v = i0 - 1;
          }
        } else if (cont) {
          s_carry.open = make_float2(tot.s, tot.c);
          s_carry.open_v = i0 - 1 + s;
        } else {
          op.item_done(i0 - 1 + s, tot.value(), s_state[s]);
        }
      };
      if (fs != ls) {
        emit(fs, ksum_cat(carry_in, first));
        if (last_ends) emit(ls, run);
      } else if (last_ends) {
        emit(ls, ksum_cat(carry_in, run));
      }
    }
    // Rows with no edges in the tile: empty rows, and a row starting exactly
    // at the tile's end (open with an empty partial).
    for (int k = 1 + tid; k <= ni; k += kRowBlock) {
      if (s_offs[k + 1] == s_offs[k]) {
        op.item_done(i0 - 1 + k, 0.0f, s_state[k]);
      } else if (s_offs[k] == j1) {
        s_carry.open = make_float2(0.f, 0.f);
        s_carry.open_v = i0 - 1 + k;
      }
    }
    __syncthreads();
    if (tid == 0) carry[t] = s_carry;
    a0 = b0;
    a1 = b1;
    b0 = c0;
    b1 = c1;
  }
  op.finish();
}

// Rows that span tiles: the tile holding a row's end folds the open partials
// of the tiles before it (from the row's first tile on) and its own tail.
template <class Op>
__global__ void __launch_bounds__(kRowBlock)
k_seg_fixup(Op op, const SegCarry *__restrict__ carry, int64_t ntiles, const int32_t *__restrict__ done) {
  if (done && *done) return;
  op.begin();
  for (int64_t t = (int64_t)blockIdx.x * kRowBlock + threadIdx.x; t < ntiles; t += (int64_t)gridDim.x * kRowBlock) {
    const int32_t v = carry[t].tail_v;
    if (v < 0) continue;
    int64_t ts = t - 1;  // carry[t - 1].open_v == v: the row was open at the previous tile's end
    while (ts > 0 && carry[ts - 1].open_v == v) ts--;
    KSum acc;
    for (int64_t q = ts; q < t; q++) acc.merge(carry[q].open.x, carry[q].open.y);
    acc.merge(carry[t].tail.x, carry[t].tail.y);
    op.item_done(v, acc.value(), op.vertex_state(v));
  }
  op.finish();
}

// One sweep: segmented rows, then the spanning-row fix-up.
template <class Op>
inline void seg_rows_launch(const Op &op, const int32_t *offs, int32_t n, int64_t m, const int32_t *parts,
                            SegCarry *carry, const int32_t *done, int grid, hipStream_t s) {
  const int64_t tiles = seg_tiles(n, m);
  hipLaunchKernelGGL(k_seg_rows<Op>, dim3(grid), dim3(kRowBlock), 0, s, op, offs, n, m, parts, carry, done);
  hipLaunchKernelGGL(k_seg_fixup<Op>, dim3(grid_for(tiles, kRowBlock, 1024)), dim3(kRowBlock), 0, s, op, carry,
                     tiles, done);
}

}  // namespace GASengine
