// Device sort and scan for the one-off graph preparation (CSC transpose,
// degree renumbering, XCD slices): a stable LSD radix sort of (key, value)
// pairs and an exclusive scan of int32 counts.  The role of the thrust
// stable sort the reference builds its CSC with (util.cuh:55-126, called
// csr_problem.cuh:754-765).
//
// Radix sort, 8 bits per pass over [begin_bit, end_bit):
//   k_radix_hist     per 2048-key tile, the 256 digit counts (LDS atomics),
//                    stored digit-major: hist[d * tiles + t]
//   scan             exclusive over the whole matrix, so entry (d, t) is
//                    where tile t's keys with digit d start
//   k_radix_scatter  each tile again, in key order: 256 keys per round, one
//                    per thread; a key's rank among equal digits below it in
//                    its wave comes from 9 ballots (its digit's bits), the
//                    waves before it from per-wave digit counts in LDS, the
//                    rounds before it from a running per-digit base, so equal
//                    keys keep their input order (stable)
// Scan: per 4096-element chunk a workgroup scan and the chunk total; one
// workgroup scans the totals; a third pass adds them.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "util.h"

namespace GASengine {

constexpr int kSortBlock = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortBlock * kSortItems;  // keys per tile
constexpr int kRadix = 256;                         // 8-bit digits
constexpr int kSortGridMax = 4096;

template <class K>
__device__ __forceinline__ int32_t radix_digit(K key, int shift) {
  return (int32_t)((key >> shift) & (K)(kRadix - 1));
}

template <class K>
__global__ void __launch_bounds__(kSortBlock)
k_radix_hist(const K *__restrict__ keys, int64_t count, int shift, int32_t *__restrict__ hist, int32_t tiles) {
  __shared__ int32_t h[kRadix];
  for (int64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
    h[threadIdx.x] = 0;  // kSortBlock == kRadix
    __syncthreads();
    const int64_t base = t * kSortTile;
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
      const int64_t i = base + r * kSortBlock + threadIdx.x;
      if (i < count) atomicAdd(&h[radix_digit(keys[i], shift)], 1);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * tiles + t] = h[threadIdx.x];
    __syncthreads();
  }
}

template <class K, class V>
__global__ void __launch_bounds__(kSortBlock)
k_radix_scatter(const K *__restrict__ keys_in, const V *__restrict__ vals_in, K *__restrict__ keys_out,
                V *__restrict__ vals_out, int64_t count, int shift, const int32_t *__restrict__ start,
                int32_t tiles) {
  constexpr int W = kSortBlock / kWave;
  __shared__ int32_t base[kRadix];
  __shared__ int32_t wcount[W][kRadix];
  const int lane = lane_id(), w = threadIdx.x / kWave;
  const uint64_t below = (1ull << lane) - 1ull;
  for (int64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
    base[threadIdx.x] = start[(int64_t)threadIdx.x * tiles + t];
    for (int r = 0; r < kSortItems; r++) {
      for (int k = threadIdx.x; k < W * kRadix; k += kSortBlock) (&wcount[0][0])[k] = 0;
      __syncthreads();
      const int64_t i = t * kSortTile + (int64_t)r * kSortBlock + threadIdx.x;
      const bool valid = i < count;
      K key = K();
      V val = V();
      int32_t d = kRadix;  // past every digit: invalid slots rank among themselves only
      if (valid) {
        key = keys_in[i];
        val = vals_in[i];
        d = radix_digit(key, shift);
      }
      uint64_t peers = ~0ull;  // lanes of this wave with the same digit
#pragma unroll
      for (int b = 0; b < 9; b++) {
        const bool bit = (d >> b) & 1;
        const uint64_t bal = ballot(bit);
        peers &= bit ? bal : ~bal;
      }
      const int32_t rank = __popcll(peers & below);
      if (valid && rank == 0) wcount[w][d] = __popcll(peers);
      __syncthreads();
      if (valid) {
        int32_t pos = base[d] + rank;
        for (int v = 0; v < w; v++) pos += wcount[v][d];
        keys_out[pos] = key;
        vals_out[pos] = val;
      }
      __syncthreads();
      {
        int32_t add = 0;
#pragma unroll
        for (int v = 0; v < W; v++) add += wcount[v][threadIdx.x];
        base[threadIdx.x] += add;
      }
    }
    __syncthreads();
  }
}

// ---- exclusive scan of int32 -------------------------------------------

constexpr int kScanChunkBlock = 1024;
constexpr int kScanChunkItems = 4;
constexpr int kScanChunk = kScanChunkBlock * kScanChunkItems;

template <int DUMMY = 0>
__global__ void __launch_bounds__(kScanChunkBlock)
k_scan_chunks(const int32_t *__restrict__ in, int32_t *__restrict__ out, int64_t count, int32_t *__restrict__ totals) {
  __shared__ int32_t smem[kScanChunkBlock / kWave + 1];
  const int64_t c = blockIdx.x;
  const int64_t first = c * kScanChunk + (int64_t)threadIdx.x * kScanChunkItems;
  int32_t v[kScanChunkItems], sum = 0;
#pragma unroll
  for (int k = 0; k < kScanChunkItems; k++) {
    v[k] = first + k < count ? in[first + k] : 0;
    sum += v[k];
  }
  int32_t total;
  int32_t ex = block_exclusive_scan<kScanChunkBlock>(sum, smem, &total);
#pragma unroll
  for (int k = 0; k < kScanChunkItems; k++) {
    if (first + k < count) out[first + k] = ex;
    ex += v[k];
  }
  if (threadIdx.x == 0) totals[c] = total;
}

// One workgroup: exclusive scan of the chunk totals in place.
template <int DUMMY = 0>
__global__ void __launch_bounds__(kScanChunkBlock) k_scan_totals(int32_t *__restrict__ totals, int64_t chunks) {
  __shared__ int32_t smem[kScanChunkBlock / kWave + 1];
  int32_t carry = 0;
  for (int64_t b = 0; b < chunks; b += kScanChunkBlock) {
    const int64_t i = b + threadIdx.x;
    const int32_t v = i < chunks ? totals[i] : 0;
    int32_t tot;
    const int32_t ex = block_exclusive_scan<kScanChunkBlock>(v, smem, &tot);
    if (i < chunks) totals[i] = carry + ex;
    carry += tot;
  }
}

template <int DUMMY = 0>
__global__ void k_scan_add(int32_t *__restrict__ out, int64_t count, const int32_t *__restrict__ totals) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    out[i] += totals[i / kScanChunk];
}

// out[i] = in[0] + ... + in[i - 1] (in and out may be the same array);
// totals: scratch of at least scan_totals_entries(count) int32.
inline int64_t scan_totals_entries(int64_t count) { return (count + kScanChunk - 1) / kScanChunk + 1; }

inline hipError_t exclusive_scan_i32(const int32_t *in, int32_t *out, int64_t count, int32_t *totals,
                                     hipStream_t s) {
  if (count <= 0) return hipSuccess;
  const int64_t chunks = (count + kScanChunk - 1) / kScanChunk;
  hipLaunchKernelGGL(k_scan_chunks<0>, dim3((unsigned)chunks), dim3(kScanChunkBlock), 0, s, in, out, count, totals);
  hipLaunchKernelGGL(k_scan_totals<0>, dim3(1), dim3(kScanChunkBlock), 0, s, totals, chunks);
  hipLaunchKernelGGL(k_scan_add<0>, dim3(grid_for(count, 256, 16384)), dim3(256), 0, s, out, count, totals);
  return hipGetLastError();
}

// Stable sort of count (key, value) pairs by key bits [begin_bit, end_bit)
// into keys_out / vals_out; the inputs are left as they are.  Temporaries
// (a second key and value buffer, the digit matrix, scan totals) are
// allocated here and freed before return; the stream is synchronised.
template <class K, class V>
hipError_t radix_sort_pairs(const K *keys_in, K *keys_out, const V *vals_in, V *vals_out, int64_t count,
                            int begin_bit, int end_bit, hipStream_t s) {
  static_assert(kSortBlock == kRadix, "one histogram bin per thread");
  if (count <= 0) return hipSuccess;
  if (count > INT32_MAX) return hipErrorInvalidValue;
  const int passes = end_bit > begin_bit ? (end_bit - begin_bit + 7) / 8 : 0;
  if (passes == 0) {
    hipError_t e = hipMemcpyAsync(keys_out, keys_in, sizeof(K) * (size_t)count, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(vals_out, vals_in, sizeof(V) * (size_t)count, hipMemcpyDeviceToDevice, s);
    return e == hipSuccess ? hipStreamSynchronize(s) : e;
  }
  const int32_t tiles = (int32_t)((count + kSortTile - 1) / kSortTile);
  const int64_t cells = (int64_t)kRadix * tiles;
  K *kt = nullptr;
  V *vt = nullptr;
  int32_t *hist = nullptr, *totals = nullptr;
  hipError_t e = hipMalloc((void **)&kt, sizeof(K) * (size_t)count);
  if (e == hipSuccess) e = hipMalloc((void **)&vt, sizeof(V) * (size_t)count);
  if (e == hipSuccess) e = hipMalloc((void **)&hist, sizeof(int32_t) * (size_t)cells);
  if (e == hipSuccess) e = hipMalloc((void **)&totals, sizeof(int32_t) * (size_t)scan_totals_entries(cells));
  if (e == hipSuccess) {
    const int grid = (int)(tiles < kSortGridMax ? tiles : kSortGridMax);
    // The last pass lands in keys_out: with an odd number of passes the
    // first one writes there too, with an even number the first writes to
    // the temporaries.
    const K *ksrc = keys_in;
    const V *vsrc = vals_in;
    for (int p = 0; p < passes && e == hipSuccess; p++) {
      const bool to_out = ((passes - 1 - p) % 2) == 0;
      K *kdst = to_out ? keys_out : kt;
      V *vdst = to_out ? vals_out : vt;
      const int shift = begin_bit + 8 * p;
      hipLaunchKernelGGL(k_radix_hist<K>, dim3(grid), dim3(kSortBlock), 0, s, ksrc, count, shift, hist, tiles);
      e = hipGetLastError();
      if (e == hipSuccess) e = exclusive_scan_i32(hist, hist, cells, totals, s);
      if (e == hipSuccess) {
        hipLaunchKernelGGL((k_radix_scatter<K, V>), dim3(grid), dim3(kSortBlock), 0, s, ksrc, vsrc, kdst, vdst, count,
                           shift, hist, tiles);
        e = hipGetLastError();
      }
      ksrc = kdst;
      vsrc = vdst;
    }
  }
  const hipError_t sy = hipStreamSynchronize(s);
  if (e == hipSuccess) e = sy;
  for (void *p : {(void *)kt, (void *)vt, (void *)hist, (void *)totals})
    if (p) (void)hipFree(p);
  return e;
}

}  // namespace GASengine
