// Compensated FP32 row-sum building blocks (KSum, strided_sum, wave_merge)
// shared by the PageRank sweeps (segreduce.h, xcd_slices.h) and the batched
// PPR kernels, and the degree bins (small / mid / large rows) used by the
// per-row PageRank set-up and gradient kernels.  Sums run in a fixed order
// with Neumaier-compensated FP32 accumulation, so results are reproducible
// run to run and accurate to about one ulp of the sum.  The reference
// materialises (vertex, value) pairs and reduces them with ReduceByKey
// (enactor_vertex_centric.cuh:2104-2424); these never write per-edge values.
//
// RowOp concept (device, by value):
//   void    begin()                       per-thread setup at kernel start
//   int32_t row_begin(int32_t v), row_end(int32_t v)
//   float   edge(int32_t v, int32_t k)    contribution of slot k of v's row
//   void    item_done(int32_t v, float)   once per vertex
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

}  // namespace GASengine
