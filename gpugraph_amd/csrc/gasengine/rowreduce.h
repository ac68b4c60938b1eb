// Row reductions over degree-binned vertex lists (PageRank forward over CSC
// in-edges, its adjoint over CSR out-edges).  Each vertex's row is reduced by
// one thread (small rows), one wave (mid) or one workgroup (large), always in
// the same order, with Neumaier-compensated FP32 accumulation, so results
// are reproducible run to run and accurate to about one ulp of the sum.
// The reference materialises (vertex, value) pairs and reduces them with
// ReduceByKey (enactor_vertex_centric.cuh:2104-2424); this never writes the
// per-edge values.
//
// RowOp concept (device, by value):
//   void    begin()                       per-thread setup at kernel start
//   int32_t row_begin(int32_t v), row_end(int32_t v)
//   float   edge(int32_t v, int32_t k)    contribution of slot k of v's row
//   void    item_done(int32_t v, float)   one thread per vertex
//   void    finish()                      all threads, once (block flushes)
#pragma once

#include "util.h"

namespace GASengine {

constexpr int kRowSmallMax = 32;    // rows up to this length: one thread
constexpr int kRowWaveMax = 4096;   // up to this: one wave; longer: one workgroup
constexpr int kRowBlock = 256;

// Neumaier-compensated FP32 sum (value = s + c).
struct KSum {
  float s = 0.f, c = 0.f;
  __device__ __forceinline__ void add(float x) {
    const float t = s + x;
    c += fabsf(s) >= fabsf(x) ? (s - t) + x : (x - t) + s;
    s = t;
  }
  __device__ __forceinline__ void merge(float s2, float c2) {
    add(s2);
    c += c2;
  }
  __device__ __forceinline__ float value() const { return s + c; }
};

// acc += edge(v, k) for k = first, first + stride, ... < end, in that order,
// four edges' loads issued before their adds (the order, and so the result,
// is that of the plain loop).
template <class Op>
__device__ __forceinline__ void strided_sum(Op &op, int32_t v, int32_t first, int32_t end, int32_t stride, KSum &acc) {
  int32_t k = first;
  for (; k + 3 * stride < end; k += 4 * stride) {
    const float x0 = op.edge(v, k), x1 = op.edge(v, k + stride);
    const float x2 = op.edge(v, k + 2 * stride), x3 = op.edge(v, k + 3 * stride);
    acc.add(x0);
    acc.add(x1);
    acc.add(x2);
    acc.add(x3);
  }
  for (; k < end; k += stride) acc.add(op.edge(v, k));
}

// Butterfly merge across the wave; lane 0 ends with the deterministic result.
__device__ __forceinline__ void wave_merge(KSum &a) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const float s2 = __shfl_xor(a.s, off), c2 = __shfl_xor(a.c, off);
    a.merge(s2, c2);
  }
}

template <class Op>
__global__ void __launch_bounds__(kRowBlock)
k_rows_small(Op op, const int32_t *__restrict__ list, int32_t count, const int32_t *__restrict__ done) {
  if (done && *done) return;
  op.begin();
  for (int64_t i = (int64_t)blockIdx.x * kRowBlock + threadIdx.x; i < count; i += (int64_t)gridDim.x * kRowBlock) {
    const int32_t v = list[i];
    KSum acc;
    strided_sum(op, v, op.row_begin(v), op.row_end(v), 1, acc);
    op.item_done(v, acc.value());
  }
  op.finish();
}

template <class Op>
__global__ void __launch_bounds__(kRowBlock)
k_rows_wave(Op op, const int32_t *__restrict__ list, int32_t count, const int32_t *__restrict__ done) {
  if (done && *done) return;
  op.begin();
  const int64_t waves = (int64_t)gridDim.x * (kRowBlock / kWave);
  for (int64_t i = (int64_t)blockIdx.x * (kRowBlock / kWave) + threadIdx.x / kWave; i < count; i += waves) {
    const int32_t v = list[i];
    KSum acc;
    strided_sum(op, v, op.row_begin(v) + lane_id(), op.row_end(v), kWave, acc);
    wave_merge(acc);
    if (lane_id() == 0) op.item_done(v, acc.value());
  }
  op.finish();
}

template <class Op>
__global__ void __launch_bounds__(kRowBlock)
k_rows_block(Op op, const int32_t *__restrict__ list, int32_t count, const int32_t *__restrict__ done) {
  __shared__ float s_s[kRowBlock / kWave], s_c[kRowBlock / kWave];
  if (done && *done) return;
  op.begin();
  for (int64_t i = blockIdx.x; i < count; i += gridDim.x) {
    const int32_t v = list[i];
    KSum acc;
    strided_sum(op, v, op.row_begin(v) + (int32_t)threadIdx.x, op.row_end(v), kRowBlock, acc);
    wave_merge(acc);
    if (lane_id() == 0) {
      s_s[threadIdx.x / kWave] = acc.s;
      s_c[threadIdx.x / kWave] = acc.c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      KSum t;
      for (int w = 0; w < kRowBlock / kWave; w++) t.merge(s_s[w], s_c[w]);
      op.item_done(v, t.value());
    }
    __syncthreads();
  }
  op.finish();
}

// Classify vertices by row length into three lists (small / mid / large).
template <int D = 0>
__global__ void __launch_bounds__(kRowBlock)
k_build_bins(const int32_t *__restrict__ off, int32_t n, int32_t *__restrict__ small, int32_t *__restrict__ mid,
             int32_t *__restrict__ large, int32_t *__restrict__ counts) {
  for (int64_t base = (int64_t)blockIdx.x * kRowBlock; base < n; base += (int64_t)gridDim.x * kRowBlock) {
    const int64_t v = base + threadIdx.x;
    int32_t d = -1;
    if (v < n) d = off[v + 1] - off[v];
    block_append<kRowBlock>(d >= 0 && d <= kRowSmallMax, (int32_t)v, small, &counts[0]);
    block_append<kRowBlock>(d > kRowSmallMax && d <= kRowWaveMax, (int32_t)v, mid, &counts[1]);
    block_append<kRowBlock>(d > kRowWaveMax, (int32_t)v, large, &counts[2]);
  }
}

// Launch the three binned kernels for one row reduction.
template <class Op>
inline void rows_launch(const Op &op, int32_t *const lists[3], const int32_t counts[3], const int32_t *done,
                        hipStream_t s) {
  if (counts[0]) hipLaunchKernelGGL(k_rows_small<Op>, dim3(grid_for(counts[0], kRowBlock, 4096)), dim3(kRowBlock), 0, s,
                                    op, lists[0], counts[0], done);
  if (counts[1])
    hipLaunchKernelGGL(k_rows_wave<Op>, dim3(grid_for((int64_t)counts[1] * kWave, kRowBlock, 4096)), dim3(kRowBlock), 0,
                       s, op, lists[1], counts[1], done);
  if (counts[2]) hipLaunchKernelGGL(k_rows_block<Op>, dim3(counts[2] < 4096 ? counts[2] : 4096), dim3(kRowBlock), 0, s,
                                    op, lists[2], counts[2], done);
}

}  // namespace GASengine
