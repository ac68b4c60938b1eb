// CSR validation and CSR -> CSC transpose on the device.
//
// Reference: the CSC and its CSC->CSR edge-id permutation are built with a
// thrust stable sort by column of the (row-sorted) edge list
// (util.cuh:55-126, called csr_problem.cuh:754-765).  Here: one stable LSD
// radix sort of (column, edge id) pairs (sort.h), so within a column the
// source rows stay ascending exactly as the reference's stable sort leaves
// them; then row ids are gathered through the permutation and the column
// offsets found by a lower-bound search per vertex.  Device memory: four
// arrays of m int32 plus the sort's temporaries, freed on return.
#pragma once

#include <hip/hip_runtime.h>

#include "sort.h"
#include "util.h"

namespace GASengine {

// bad[0] / bad[1]: count and lowest index of row-offset violations (offsets
// not starting at 0, decreasing, or not ending at m); bad[2] / bad[3]: count
// and lowest slot of column ids outside [0, n).  Zero bad[0], bad[2] and set
// bad[1], bad[3] to INT32_MAX before the launch.
template <int DUMMY = 0>
__global__ void k_validate_csr(const int32_t *__restrict__ ro, const int32_t *__restrict__ col, int32_t n, int64_t m,
                               int32_t *__restrict__ bad) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int64_t v = t0; v <= n; v += stride) {
    const bool ok = v == 0 ? ro[0] == 0 : (ro[v - 1] <= ro[v] && (v < n || ro[n] == m));
    if (!ok) {
      atomicAdd(&bad[0], 1);
      atomicMin(&bad[1], (int32_t)v);
    }
  }
  for (int64_t e = t0; e < m; e += stride) {
    const int32_t c = col[e];
    if (c < 0 || c >= n) {
      atomicAdd(&bad[2], 1);
      atomicMin(&bad[3], (int32_t)(e < INT32_MAX ? e : INT32_MAX));
    }
  }
}

// rows[e] = the row owning slot e (one wave per row).
template <int DUMMY = 0>
__global__ void k_edge_rows(const int32_t *__restrict__ ro, int32_t n, int32_t *__restrict__ rows) {
  const int64_t waves = (int64_t)gridDim.x * (blockDim.x / kWave);
  for (int64_t u = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; u < n; u += waves)
    for (int32_t e = ro[u] + lane_id(); e < ro[u + 1]; e += kWave) rows[e] = (int32_t)u;
}

template <int DUMMY = 0>
__global__ void k_iota(int32_t *__restrict__ p, int64_t m) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = (int32_t)i;
}

template <int DUMMY = 0>
__global__ void k_gather_rows(const int32_t *__restrict__ rows, const int32_t *__restrict__ eid, int64_t m,
                              int32_t *__restrict__ ri) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += (int64_t)gridDim.x * blockDim.x)
    ri[k] = rows[eid[k]];
}

// co[v] = first position of a column >= v in the sorted keys (v = 0..n).
template <int DUMMY = 0>
__global__ void k_column_offsets(const int32_t *__restrict__ keys, int64_t m, int32_t n, int32_t *__restrict__ co) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v <= n; v += (int64_t)gridDim.x * blockDim.x) {
    int64_t lo = 0, hi = m;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (keys[mid] < v)
        lo = mid + 1;
      else
        hi = mid;
    }
    co[v] = (int32_t)lo;
  }
}

// Transpose a validated device CSR into caller-provided co[n+1], ri[m],
// eid[m] on stream s (synchronises s; temporaries are freed before return).
// Returns the first HIP error, allocations included.
inline hipError_t device_csr_to_csc(int32_t n, int64_t m, const int32_t *ro, const int32_t *col, int32_t *co,
                                    int32_t *ri, int32_t *eid, hipStream_t s) {
  if (m == 0) {
    hipLaunchKernelGGL(k_fill<int32_t>, dim3(grid_for((int64_t)n + 1, 256, 4096)), dim3(256), 0, s, co, 0,
                       (int64_t)n + 1);
    return hipGetLastError();
  }
  int32_t *keys_out = nullptr, *ids = nullptr, *rows = nullptr;
  int end_bit = 1;
  while (end_bit < 31 && ((int64_t)1 << end_bit) < n) end_bit++;
  hipError_t e = hipMalloc((void **)&keys_out, 4 * (size_t)m);
  if (e == hipSuccess) e = hipMalloc((void **)&ids, 4 * (size_t)m);
  if (e == hipSuccess) e = hipMalloc((void **)&rows, 4 * (size_t)m);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_iota<0>, dim3(grid_for(m, 256, 16384)), dim3(256), 0, s, ids, m);
    hipLaunchKernelGGL(k_edge_rows<0>, dim3(grid_for((int64_t)n * kWave, 256, 16384)), dim3(256), 0, s, ro, n, rows);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = radix_sort_pairs<int32_t, int32_t>(col, keys_out, ids, eid, m, 0, end_bit, s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_gather_rows<0>, dim3(grid_for(m, 256, 16384)), dim3(256), 0, s, rows, eid, m, ri);
    hipLaunchKernelGGL(k_column_offsets<0>, dim3(grid_for((int64_t)n + 1, 256, 16384)), dim3(256), 0, s, keys_out, m,
                       n, co);
    e = hipGetLastError();
  }
  const hipError_t sy = hipStreamSynchronize(s);
  if (e == hipSuccess) e = sy;
  for (void *p : {(void *)keys_out, (void *)ids, (void *)rows})
    if (p) (void)hipFree(p);
  return e;
}

}  // namespace GASengine
