// XCD-sliced row reductions (PageRank sweeps over graphs whose gathered
// vector is larger than one L2).
//
// gfx950 has 8 XCDs, each with a private 4 MB L2, and dispatches consecutive
// workgroup ids round-robin over them: workgroup b runs on XCD b % 8.  A row
// reduction sum_k x[idx[k]] over a vector of 4M floats (16 MB) misses L2 on
// almost every gather.  Here each row's edge list is split by the slice of the
// gathered vertex, slice(u) = (u >> 5) & 7 -- 32 consecutive vertices, one
// 128-byte line of x, per slice, so the slices interleave line by line and
// stay balanced under any vertex ordering -- and workgroup b reduces only the
// sub-rows of slice b % 8.  Each XCD's L2 then serves one eighth of x.  Every
// (slice, row) reduction writes a Neumaier-compensated partial (s, c); the
// combine pass folds a row's 8 partials in slice order, so the result is
// reproducible run to run and within a few ulp of the exact sum.
//
// SliceSet (per graph and direction), built once on the device:
//   off[s * n + v], v = 0..n   offsets of slice s's sub-rows into idx / val
//                              (slices stored one after another: off[s*n+n] is
//                              slice s+1's first offset; off[8n] = m)
//   idx[k]                     gathered vertex of slice-ordered slot k
//   perm[k]                    the original slot (to gather edge values)
//   mid / large                rows whose slice length exceeds kRowSmallMax /
//                              kRowWaveMax, as s * n + v, grouped by slice
//                              (host bases mid_base[9], large_base[9])
//
// RowOp concept: as rowreduce.h (begin, edge(v, k), item_done(v, sum),
// finish); edge() indexes the slice-ordered arrays.
#pragma once

#include "rowreduce.h"

namespace GASengine {

constexpr int kXcdSlices = 8;

__host__ __device__ __forceinline__ int xcd_slice_of(int32_t u) { return (u >> 5) & (kXcdSlices - 1); }

// ---- build ---------------------------------------------------------------

// cnt[s * n + v] = edges of row v whose gathered vertex is in slice s.  One
// wave per row.
template <int DUMMY = 0>
__global__ void __launch_bounds__(256)
k_slice_count(const int32_t *__restrict__ off, const int32_t *__restrict__ nbr, int32_t n, int32_t *__restrict__ cnt) {
  const int64_t waves = (int64_t)gridDim.x * (256 / kWave);
  for (int64_t v = (int64_t)blockIdx.x * (256 / kWave) + threadIdx.x / kWave; v < n; v += waves) {
    int32_t c[kXcdSlices] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int32_t b = off[v], e = off[v + 1];
    for (int32_t k = b + lane_id(); k < e; k += kWave) {
      const int s = xcd_slice_of(nbr[k]);
#pragma unroll
      for (int t = 0; t < kXcdSlices; t++) c[t] += s == t;
    }
#pragma unroll
    for (int t = 0; t < kXcdSlices; t++) {
      const int32_t tot = wave_sum(c[t]);
      if (lane_id() == 0) cnt[(int64_t)t * n + v] = tot;
    }
  }
}

// Stable scatter of each row's slots into the slice-ordered arrays.  One wave
// per row, 64 slots per step; per slice a wave-uniform running position.
template <int DUMMY = 0>
__global__ void __launch_bounds__(256)
k_slice_fill(const int32_t *__restrict__ off, const int32_t *__restrict__ nbr, int32_t n,
             const int32_t *__restrict__ soff, int32_t *__restrict__ idx, int32_t *__restrict__ perm) {
  const int64_t waves = (int64_t)gridDim.x * (256 / kWave);
  const int lane = lane_id();
  for (int64_t v = (int64_t)blockIdx.x * (256 / kWave) + threadIdx.x / kWave; v < n; v += waves) {
    int32_t pos[kXcdSlices];
#pragma unroll
    for (int t = 0; t < kXcdSlices; t++) pos[t] = soff[(int64_t)t * n + v];
    const int32_t b = off[v], e = off[v + 1];
    for (int32_t k0 = b; k0 < e; k0 += kWave) {
      const int32_t k = k0 + lane;
      const bool ok = k < e;
      const int32_t u = ok ? nbr[k] : 0;
      const int s = ok ? xcd_slice_of(u) : -1;
      int32_t mine = 0;
#pragma unroll
      for (int t = 0; t < kXcdSlices; t++) {
        const uint64_t mask = ballot(s == t);
        if (s == t) mine = pos[t] + rank_below(mask);
        pos[t] += __popcll(mask);
      }
      if (ok) {
        idx[mine] = u;
        perm[mine] = k;
      }
    }
  }
}

// Long sub-rows: pass 0 counts per (slice, class), pass 1 appends s * n + v
// at base[class][s] + running count.  ctr: [2][8] counters.
template <int PASS>
__global__ void __launch_bounds__(256)
k_slice_long(const int32_t *__restrict__ soff, int32_t n, int32_t *__restrict__ ctr, const int32_t *__restrict__ base,
             int32_t *__restrict__ mid, int32_t *__restrict__ large) {
  const int64_t total = (int64_t)kXcdSlices * n;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t len = soff[i + 1] - soff[i];
    if (len <= kRowSmallMax) continue;
    const int s = (int)(i / n);
    const int cls = len <= kRowWaveMax ? 0 : 1;
    const int32_t at = atomicAdd(&ctr[cls * kXcdSlices + s], 1);
    if (PASS == 1) (cls == 0 ? mid : large)[base[cls * (kXcdSlices + 1) + s] + at] = (int32_t)i;
  }
}

template <class T>
__global__ void k_gather_perm(const T *__restrict__ in, const int32_t *__restrict__ perm, int64_t m, T *__restrict__ out) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += (int64_t)gridDim.x * blockDim.x)
    out[k] = in[perm[k]];
}

// ---- sweep ---------------------------------------------------------------

struct SliceLists {
  const int32_t *off;    // [8n + 1]
  const int32_t *mid;    // grouped by slice
  const int32_t *large;
  int32_t n;
  int32_t mid_base[kXcdSlices + 1];
  int32_t large_base[kXcdSlices + 1];
};

// Short sub-rows, one thread each; workgroup b covers slice b % 8.
template <class Op>
__global__ void __launch_bounds__(kRowBlock)
k_slice_small(Op op, SliceLists L, float2 *__restrict__ part, const int32_t *__restrict__ done) {
  if (done && *done) return;
  op.begin();
  const int s = blockIdx.x % kXcdSlices;
  const int64_t per = gridDim.x / kXcdSlices;
  const int32_t *so = L.off + (int64_t)s * L.n;
  float2 *ps = part + (int64_t)s * L.n;
  for (int64_t v = (int64_t)(blockIdx.x / kXcdSlices) * kRowBlock + threadIdx.x; v < L.n; v += per * kRowBlock) {
    const int32_t b = so[v], e = so[v + 1];
    if (e - b > kRowSmallMax) continue;
    KSum acc;
    strided_sum(op, (int32_t)v, b, e, 1, acc);
    ps[v] = make_float2(acc.s, acc.c);
  }
}

// Mid sub-rows, one wave each.
template <class Op>
__global__ void __launch_bounds__(kRowBlock)
k_slice_mid(Op op, SliceLists L, float2 *__restrict__ part, const int32_t *__restrict__ done) {
  if (done && *done) return;
  op.begin();
  const int s = blockIdx.x % kXcdSlices;
  const int64_t per = gridDim.x / kXcdSlices;
  constexpr int W = kRowBlock / kWave;
  const int32_t lo = L.mid_base[s], hi = L.mid_base[s + 1];
  for (int64_t i = lo + (int64_t)(blockIdx.x / kXcdSlices) * W + threadIdx.x / kWave; i < hi; i += per * W) {
    const int32_t sv = L.mid[i];
    const int32_t v = sv - s * L.n;
    const int32_t b = L.off[sv], e = L.off[sv + 1];
    KSum acc;
    strided_sum(op, v, b + lane_id(), e, kWave, acc);
    wave_merge(acc);
    if (lane_id() == 0) part[sv] = make_float2(acc.s, acc.c);
  }
}

// Large sub-rows, one workgroup each.
template <class Op>
__global__ void __launch_bounds__(kRowBlock)
k_slice_large(Op op, SliceLists L, float2 *__restrict__ part, const int32_t *__restrict__ done) {
  __shared__ float s_s[kRowBlock / kWave], s_c[kRowBlock / kWave];
  if (done && *done) return;
  op.begin();
  const int s = blockIdx.x % kXcdSlices;
  const int64_t per = gridDim.x / kXcdSlices;
  const int32_t lo = L.large_base[s], hi = L.large_base[s + 1];
  for (int64_t i = lo + blockIdx.x / kXcdSlices; i < hi; i += per) {
    const int32_t sv = L.large[i];
    const int32_t v = sv - s * L.n;
    const int32_t b = L.off[sv], e = L.off[sv + 1];
    KSum acc;
    strided_sum(op, v, b + (int32_t)threadIdx.x, e, kRowBlock, acc);
    wave_merge(acc);
    if (lane_id() == 0) {
      s_s[threadIdx.x / kWave] = acc.s;
      s_c[threadIdx.x / kWave] = acc.c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      KSum t;
      for (int w = 0; w < kRowBlock / kWave; w++) t.merge(s_s[w], s_c[w]);
      part[sv] = make_float2(t.s, t.c);
    }
    __syncthreads();
  }
}

// Fold the 8 partials of each row in slice order and hand the sum to the op.
template <class Op>
__global__ void __launch_bounds__(kRowBlock)
k_slice_combine(Op op, int32_t n, const float2 *__restrict__ part, const int32_t *__restrict__ done) {
  if (done && *done) return;
  op.begin();
  for (int64_t v = (int64_t)blockIdx.x * kRowBlock + threadIdx.x; v < n; v += (int64_t)gridDim.x * kRowBlock) {
    KSum acc;
#pragma unroll
    for (int s = 0; s < kXcdSlices; s++) {
      const float2 p = part[(int64_t)s * n + v];
      acc.merge(p.x, p.y);
    }
    op.item_done((int32_t)v, acc.value());
  }
  op.finish();
}

// One sweep: the three partial kernels (grids multiples of 8) then the combine.
template <class Op>
inline void slice_rows_launch(const Op &op, const SliceLists &L, float2 *part, const int32_t *done, hipStream_t s) {
  const int per_small = grid_for(L.n, kRowBlock, 512);
  hipLaunchKernelGGL(k_slice_small<Op>, dim3(kXcdSlices * per_small), dim3(kRowBlock), 0, s, op, L, part, done);
  int32_t mid_max = 0, large_max = 0;
  for (int t = 0; t < kXcdSlices; t++) {
    mid_max = mid_max > L.mid_base[t + 1] - L.mid_base[t] ? mid_max : L.mid_base[t + 1] - L.mid_base[t];
    large_max = large_max > L.large_base[t + 1] - L.large_base[t] ? large_max : L.large_base[t + 1] - L.large_base[t];
  }
  if (mid_max)
    hipLaunchKernelGGL(k_slice_mid<Op>, dim3(kXcdSlices * grid_for((int64_t)mid_max * kWave, kRowBlock, 512)),
                       dim3(kRowBlock), 0, s, op, L, part, done);
  if (large_max)
    hipLaunchKernelGGL(k_slice_large<Op>, dim3(kXcdSlices * (large_max < 512 ? large_max : 512)), dim3(kRowBlock), 0, s,
                       op, L, part, done);
  hipLaunchKernelGGL(k_slice_combine<Op>, dim3(grid_for(L.n, kRowBlock, 4096)), dim3(kRowBlock), 0, s, op, L.n, part,
                     done);
}

}  // namespace GASengine
