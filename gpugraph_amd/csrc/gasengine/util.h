// Device primitives for the GAS engine on CDNA (wave64): lane masks, wave
// aggregated queue append, workgroup scans, bitmap helpers, fill kernels.
// Replaces the roles of b40c/util (scan, tile I/O, memset/sequence kernels)
// and moderngpu's CTA scan for this engine.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <initializer_list>
#include <utility>

namespace GASengine {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// Number of set bits of `mask` strictly below this lane.
__device__ __forceinline__ int rank_below(uint64_t mask) {
  return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                        __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Append `value` to queue[*size] for every lane with keep set: one atomic per
// wave (leader reserves popcount slots), order within the wave by lane.
// Must be reached by all lanes of the wave.
__device__ __forceinline__ void wave_append(bool keep, int32_t value, int32_t *queue, int32_t *size) {
  uint64_t mask = ballot(keep);
  if (mask == 0) return;
  int leader = __builtin_ctzll(mask);
  int base = 0;
  if (lane_id() == leader) base = atomicAdd(size, (int32_t)__popcll(mask));
  base = __shfl(base, leader);
  if (keep) queue[base + rank_below(mask)] = value;
}

// Set bit v of bits for every lane with `want`.  Rounds combine the lanes
// whose bits share a word (the lowest lane's word each round) into one
// request by that lane; they stop after kMarkRounds, or after two rounds
// that found a lane alone, since scattered bits gain nothing from combining.
// Each combined word, and each bit left over, is read first (one load per
// lane, all issued together) and set with an atomicOr only where bits are
// missing: bits already set, such as a hub's marked by an earlier wave, cost
// no atomic.  The atomics' results are unused, so the wave does not wait for
// them.  Must be reached by all lanes of the wave.
constexpr int kMarkRounds = 8;
__device__ __forceinline__ void wave_mark(bool want, int32_t v, uint32_t *bits) {
  uint64_t todo = ballot(want);
  if (todo == 0) return;
  const int lane = lane_id();
  const int32_t w = v >> 5;
  const uint32_t m = 1u << (v & 31);
  int32_t my_word = -1;  // the word this lane sets, and its bits
  uint32_t my_bits = 0u;
  int alone = 0;
  for (int r = 0; r < kMarkRounds && todo; r++) {
    const int leader = __builtin_ctzll(todo);
    const int32_t lw = __builtin_amdgcn_readlane(w, leader);
    const bool peer = ((todo >> lane) & 1ull) && w == lw;
    const uint64_t peers = ballot(peer);
    if (__popcll(peers) == 1 && ++alone == 2) break;  // the leader stays in todo and marks on its own
    uint32_t acc = peer ? m : 0u;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc |= __shfl_xor(acc, off);
    if (lane == leader) {
      my_word = lw;
      my_bits = acc;
    }
    todo &= ~peers;
  }
  if ((todo >> lane) & 1ull) {
    my_word = w;
    my_bits = m;
  }
  if (my_word >= 0) {
    const uint32_t cur = bits[my_word];
    if ((cur & my_bits) != my_bits) atomicOr(&bits[my_word], my_bits);
  }
}

// Wave-wide reductions (all lanes participate).
template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Inclusive scan across the wave.
__device__ __forceinline__ int32_t wave_inclusive_scan(int32_t v) {
  const int l = lane_id();
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    int32_t t = __shfl_up(v, off);
    if (l >= off) v += t;
  }
  return v;
}

// Exclusive scan over a workgroup of BLOCK threads.  `smem` needs
// BLOCK/64 + 1 ints.  Returns the exclusive prefix; *total gets the sum.
template <int BLOCK>
__device__ __forceinline__ int32_t block_exclusive_scan(int32_t v, int32_t *smem, int32_t *total) {
  constexpr int W = BLOCK / kWave;
  const int wid = threadIdx.x / kWave, l = lane_id();
  int32_t inc = wave_inclusive_scan(v);
  if (l == kWave - 1) smem[wid] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t run = 0;
#pragma unroll
    for (int w = 0; w < W; w++) {
      int32_t t = smem[w];
      smem[w] = run;
      run += t;
    }
    smem[W] = run;
  }
  __syncthreads();
  int32_t excl = smem[wid] + inc - v;
  *total = smem[W];
  __syncthreads();
  return excl;
}

// Workgroup sum of 64-bit values; result valid in every thread.
template <int BLOCK>
__device__ __forceinline__ unsigned long long block_sum_u64(unsigned long long v, unsigned long long *smem) {
  constexpr int W = BLOCK / kWave;
  v = wave_sum(v);
  const int wid = threadIdx.x / kWave;
  if (lane_id() == 0) smem[wid] = v;
  __syncthreads();
  unsigned long long t = 0;
#pragma unroll
  for (int w = 0; w < W; w++) t += smem[w];
  __syncthreads();
  return t;
}

// Striped counters.  Device-scope atomics are performed beyond the per-XCD
// L2s, and same-address ones from many workgroups serialise there: on
// MI355X one atomic per workgroup of a 2048-workgroup launch into a single
// counter added about 30 us to the launch (bottom-up BFS: 131 -> 55 us with
// its four counters striped).  A counter that every workgroup adds to is
// therefore kStripes words, STRIDE words apart (at least 128 bytes):
// workgroup b adds to word (b % kStripes) * STRIDE, and readers fold the
// stripes (stripe_sum, stripe_min, or the host-side read of a graph's
// counter array).  A graph's counter array holds kCounterSlots counters per
// stripe.
constexpr int kStripes = 16;
constexpr int kCounterSlots = 64;
constexpr int kCounterWords = kStripes * kCounterSlots;

template <int STRIDE = kCounterSlots, class T>
__device__ __forceinline__ T *stripe_of(T *counter) {
  return counter + (int)(blockIdx.x % kStripes) * STRIDE;
}
template <int STRIDE = kCounterSlots, class T>
__device__ __forceinline__ T stripe_sum(const T *counter) {
  T t = counter[0];
#pragma unroll
  for (int q = 1; q < kStripes; q++) t += counter[q * STRIDE];
  return t;
}

// Add a per-thread value to a striped counter with one atomic per
// workgroup.  All threads of the workgroup must call it.
template <int BLOCK, int STRIDE = kCounterSlots>
__device__ __forceinline__ void block_counter_add(unsigned long long *counter, unsigned long long v) {
  __shared__ unsigned long long red[BLOCK / kWave];
  const unsigned long long t = block_sum_u64<BLOCK>(v, red);
  if (threadIdx.x == 0 && t) atomicAdd(stripe_of<STRIDE>(counter), t);
}

// ---- single-pass scan across workgroups ------------------------------
//
// For grids of at most BLOCK workgroups: workgroup b publishes its aggregate
// in status[b] (bit 63 set over a 63-bit value), then its threads read the
// words of workgroups 0..b-1 at once, one each, waiting on any not yet
// published, and sum them.  No word is written twice and no address takes a
// second write, so nothing serialises the way one shared atomicAdd per
// workgroup does, and there is no chain of prefixes to wait along.  Values
// are sums of packed fields that must not carry into each other (the caller's
// packing keeps every field's total in range).  The words start at 0 for
// every scan.  Workgroups wait only on lower-numbered ones, which are
// dispatched before them.  All threads of the workgroup must call it; a
// workgroup that does not need its prefix (need false, the same in all its
// threads) only publishes and gets 0.
constexpr uint64_t kScanPublished = 1ull << 63;
constexpr uint64_t kScanValueMask = kScanPublished - 1;

template <int BLOCK>
__device__ __forceinline__ uint64_t block_exclusive_across(uint64_t *status, int b, uint64_t aggregate, bool need) {
  __shared__ unsigned long long red[BLOCK / kWave];
  if (threadIdx.x == 0)
    __hip_atomic_store(&status[b], kScanPublished | aggregate, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (!need) return 0;
  unsigned long long v = 0;
  if ((int)threadIdx.x < b) {
    uint64_t w;
    while (!((w = __hip_atomic_load(&status[threadIdx.x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) &
             kScanPublished))
      __builtin_amdgcn_s_sleep(1);
    v = w & kScanValueMask;
  }
  return block_sum_u64<BLOCK>(v, red);
}

// Workgroup-aggregated queue append: one global atomic per call.  All
// threads of the workgroup must call it.
template <int BLOCK>
__device__ __forceinline__ void block_append(bool keep, int32_t value, int32_t *queue, int32_t *size) {
  __shared__ int32_t scan[BLOCK / kWave + 1];
  __shared__ int32_t base;
  int32_t total;
  const int32_t off = block_exclusive_scan<BLOCK>(keep ? 1 : 0, scan, &total);
  if (total == 0) return;
  if (threadIdx.x == 0) base = atomicAdd(size, total);
  __syncthreads();
  if (keep) queue[base + off] = value;
  __syncthreads();
}

// a[i] = v for every i < n except a[at] = av, by the threads t0, t0 +
// stride, ...; 16-byte stores when a is 16-byte aligned (the tail and an
// unaligned array one element per store).
__device__ __forceinline__ void fill_except_one(int32_t *a, int64_t n, int32_t v, int64_t at, int32_t av,
                                                int64_t t0, int64_t stride) {
  int64_t done = 0;
  if ((reinterpret_cast<uintptr_t>(a) & 15) == 0) {
    const int64_t n4 = n >> 2;
    int4 *a4 = reinterpret_cast<int4 *>(a);
    const int64_t at4 = at >> 2;
    const int k = (int)(at & 3);
    for (int64_t i = t0; i < n4; i += stride) {
      int4 x = make_int4(v, v, v, v);
      if (i == at4) {
        x.x = k == 0 ? av : v;
        x.y = k == 1 ? av : v;
        x.z = k == 2 ? av : v;
        x.w = k == 3 ? av : v;
      }
      a4[i] = x;
    }
    done = n4 << 2;
  }
  for (int64_t i = done + t0; i < n; i += stride) a[i] = i == at ? av : v;
}

// ---- bitmaps (bit v of word v>>5) --------------------------------------

__device__ __forceinline__ bool bit_test(const uint32_t *bits, int32_t v) {
  return (bits[v >> 5] >> (v & 31)) & 1u;
}

// Atomic test-and-set; true iff this call flipped the bit from 0 to 1.  A
// plain (cacheable) read first skips the atomic for bits already set; bits
// are only set while kernels run, so a stale read can only miss a set bit,
// and the atomic's returned word then decides.
__device__ __forceinline__ bool bit_test_and_set(uint32_t *bits, int32_t v) {
  const uint32_t m = 1u << (v & 31);
  if (bits[v >> 5] & m) return false;
  return (atomicOr(&bits[v >> 5], m) & m) == 0;
}

// ---- fills --------------------------------------------------------------

template <class T>
__global__ void k_fill(T *p, T value, int64_t n) {
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
  if constexpr (16 % sizeof(T) == 0) {
    // 16-byte stores between a head and a tail of fewer than E elements each
    // (the grid has more than E threads).
    constexpr int E = 16 / sizeof(T);
    const uintptr_t a = reinterpret_cast<uintptr_t>(p);
    int64_t head = a % sizeof(T) ? n : (int64_t)(((16 - (a & 15)) & 15) / sizeof(T));
    head = head < n ? head : n;
    const int64_t body = (n - head) / E, t0 = head + body * E;
    union {
      T t[E];
      uint4 v;
    } pat;
#pragma unroll
    for (int k = 0; k < E; k++) pat.t[k] = value;
    uint4 *q = reinterpret_cast<uint4 *>(p + head);
    for (int64_t i = tid; i < body; i += stride) q[i] = pat.v;
    for (int64_t i = tid; i < head; i += stride) p[i] = value;
    for (int64_t i = t0 + tid; i < n; i += stride) p[i] = value;
  } else {
    for (int64_t i = tid; i < n; i += stride) p[i] = value;
  }
}

template <class T>
__global__ void k_sequence(T *p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = (T)i;
}

// Several 64-bit counter slots set in one launch (per-iteration resets).
struct SlotResets {
  int32_t count;
  int32_t idx[8];
  unsigned long long val[8];
};

// Sets the listed counters of a graph's counter array in every stripe.
template <int D = 0>
__global__ void k_slot_resets(unsigned long long *c, SlotResets r) {
  const int e = threadIdx.x / kStripes, q = threadIdx.x % kStripes;
  if (e < r.count) c[q * kCounterSlots + r.idx[e]] = r.val[e];
}

inline hipError_t slot_resets_async(unsigned long long *c, std::initializer_list<std::pair<int, unsigned long long>> kv,
                                    hipStream_t s) {
  SlotResets r{};
  if (kv.size() > 8) return hipErrorInvalidValue;
  for (const auto &p : kv) {
    r.idx[r.count] = p.first;
    r.val[r.count] = p.second;
    r.count++;
  }
  hipLaunchKernelGGL(k_slot_resets<0>, dim3(1), dim3(8 * kStripes), 0, s, c, r);
  return hipGetLastError();
}

inline int grid_for(int64_t n, int block, int max_grid = 65536) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > max_grid) g = max_grid;
  return (int)g;
}

template <class T>
inline hipError_t fill_async(T *p, T value, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int64_t per = 16 % sizeof(T) == 0 ? 16 / sizeof(T) : 1;  // elements per thread store
  hipLaunchKernelGGL(k_fill<T>, dim3(grid_for((n + per - 1) / per, 256, 8192)), dim3(256), 0, s, p, value, n);
  return hipGetLastError();
}

template <class T>
inline hipError_t sequence_async(T *p, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sequence<T>, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

}  // namespace GASengine
