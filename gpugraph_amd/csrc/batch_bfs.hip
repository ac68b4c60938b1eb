// Batched multi-source BFS (gg_batch_run, num_bfs > 0): up to 64 BFS
// instances advance together, one bit per instance in 64-bit words per
// vertex, replacing the reference drivers' sequential per-source loop
// (Algorithms/BFS/bfs.cu:865-881, one EnactIterativeSearch per source).
// Result contract per instance: the reference BFS program's depths
// (Algorithms/BFS/bfs.h; CPUBFS bfs.cu:57-111): depth[src] = 0, -1 where
// unreached.
//
// State per vertex (8 bytes each): seen (instances that reached it), F
// (instances for which it is on the current frontier; two copies, current
// and next), next (instances reaching it at this level).  A level is
//   top-down   LBS over the union frontier queue (vertices with F != 0):
//              per out-edge (u, w): bits = F[u] & ~seen[w]; atomicOr into
//              next[w] when it adds a bit (plain pre-check first);
//   bottom-up  one lane per vertex not yet reached by every instance: OR
//              of F over the in-neighbours, stopping once every missing
//              instance is found; the owner writes next[v] with a plain
//              store.  In-rows longer than 32 are cut into chunks of at most
//              1024 in-edges (listed once per graph) that the waves share
//              out evenly, 64 in-edges per step, each chunk stopping early the
//              same way and folding its bits into next[v] with an atomicOr:
//              a wave per whole row left the waves holding the hubs (the
//              lowest ids after renumbering) with most of the level's work;
//   update     per vertex: next -> seen, F_next = next, the vertex joins the
//              next union queue, and depth[i][v] = level + 1 for every
//              instance bit i of next (a wave's lanes write instance i
//              together: consecutive vertices, consecutive addresses).
// Direction: bottom-up when the union frontier's out-edges exceed m / 20.
// Depths are written straight to the instance-major output (at the caller's
// vertex ids on a renumbered graph).
//
// Byte model per level: top-down 12 F + 16 E (column + F[u] + seen/next
// probe) + 8 N; bottom-up 16 U + 12 E (U vertices scanned, E in-edges
// examined); update 32 n + 4 k N.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>

#include "internal.h"

namespace gg {
namespace {

using namespace GASengine;
constexpr int kBfsBatchMax = 64;
constexpr int kMsMfSlot = 56;        // counter: out-degree sum of the next union frontier
constexpr int kMsExaminedSlot = 57;  // counter: in-edges examined by bottom-up steps

// Top-down edge operation over CSR rows: instances of F[u] that have not
// reached w join next[w].  Emits nothing: the update kernel builds the next
// union queue.
struct MsPushOp {
  static constexpr bool kItemValues = false;
  const int32_t *ro;
  const int32_t *col;
  const unsigned long long *F;
  const unsigned long long *seen;
  unsigned long long *next;

  __device__ __forceinline__ void prologue(int32_t) {}
  __device__ __forceinline__ int32_t degree(int32_t u) { return ro[u + 1] - ro[u]; }
  __device__ __forceinline__ int32_t row_begin(int32_t u) { return ro[u]; }
  __device__ __forceinline__ int32_t item_value(int32_t) { return 0; }
  __device__ __forceinline__ bool edge(int32_t u, int32_t, int32_t slot, int32_t &) {
    const int32_t w = col[slot];
    const unsigned long long bits = F[u] & ~seen[w];
    if (bits && (bits & ~next[w])) atomicOr(&next[w], bits);
    return false;
  }
  __device__ __forceinline__ void finish() {}
};

__device__ __forceinline__ unsigned long long wave_or(unsigned long long x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x |= __shfl_xor(x, o);
  return x;
}

constexpr int32_t kMsLongRow = 32;  // in-rows longer than this are scanned in chunks by whole waves
constexpr int32_t kMsChunk = 1024;  // in-edges per chunk

__global__ void __launch_bounds__(256)
k_ms_bottom_up(const int32_t *__restrict__ co, const int32_t *__restrict__ ri, const unsigned long long *__restrict__ F,
               const unsigned long long *__restrict__ seen, unsigned long long *__restrict__ next, int32_t n,
               unsigned long long full, const int4 *__restrict__ chunks, int64_t nchunks,
               unsigned long long *__restrict__ counters) {
  const int lane = lane_id();
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  unsigned long long examined = 0;
  for (int64_t p = (((int64_t)blockIdx.x * blockDim.x) + threadIdx.x) >> 6; p * 64 < n; p += nwaves) {
    const int32_t v = (int32_t)(p * 64 + lane);
    unsigned long long sv = full;
    int32_t b = 0, e = 0;
    if (v < n) {
      sv = seen[v];
      b = co[v];
      e = co[v + 1];
    }
    const bool todo = v < n && sv != full && e > b && e - b <= kMsLongRow;
    if (todo) {
      unsigned long long acc = 0;
      int32_t k = b;
      for (; k + 4 <= e && ((acc | sv) != full); k += 4) {
        const int32_t u0 = ri[k], u1 = ri[k + 1], u2 = ri[k + 2], u3 = ri[k + 3];
        acc |= F[u0] | F[u1] | F[u2] | F[u3];
        examined += 4;
      }
      for (; k < e && ((acc | sv) != full); k++) {
        acc |= F[ri[k]];
        examined++;
      }
      const unsigned long long nx = acc & ~sv;
      if (nx) next[v] = nx;
    }
  }
  // Chunks of the long in-rows (vertex, first in-edge, end): the whole wave,
  // 64 in-edges per step; chunks of one row fold into next[v] atomically.
  for (int64_t q = (((int64_t)blockIdx.x * blockDim.x) + threadIdx.x) >> 6; q < nchunks; q += nwaves) {
    const int4 ch = chunks[q];
    const unsigned long long wsv = seen[ch.x];
    if (wsv == full) continue;
    unsigned long long acc = 0;
    for (int32_t k0 = ch.y; k0 < ch.z && ((acc | wsv) != full); k0 += kWave) {
      const int32_t k = k0 + lane;
      acc |= wave_or(k < ch.z ? F[ri[k]] : 0ull);
      if (lane == 0) examined += (unsigned long long)min(kWave, ch.z - k0);
    }
    const unsigned long long nx = acc & ~wsv;
    if (lane == 0 && nx) atomicOr(&next[ch.x], nx);
  }
  block_counter_add<256>(&counters[kMsExaminedSlot], examined);
}

// The chunk list of the long in-rows, built once per graph on the host from
// the in-edge offsets: (v, b, e, 0) per chunk of at most kMsChunk in-edges.
int ms_long_chunks(gg_graph *g, const int4 **out, int64_t *count) {
  if (g->ms_chunk_count < 0) {
    const int32_t n = g->n;
    std::vector<int32_t> co((size_t)n + 1);
    GG_HIP(hipMemcpyAsync(co.data(), g->dg.column_offsets, 4 * ((size_t)n + 1), hipMemcpyDeviceToHost, g->stream));
    GG_HIP(hipStreamSynchronize(g->stream));
    std::vector<int4> ch;
    for (int32_t v = 0; v < n; v++) {
      const int32_t b = co[(size_t)v], e = co[(size_t)v + 1];
      if (e - b <= kMsLongRow) continue;
      for (int32_t k = b; k < e; k += kMsChunk) ch.push_back(make_int4(v, k, std::min(k + kMsChunk, e), 0));
    }
    GG_TRY(g->ms_chunks.ensure(sizeof(int4) * std::max<size_t>(ch.size(), 1)));
    if (!ch.empty())
      GG_HIP(hipMemcpyAsync(g->ms_chunks.p, ch.data(), sizeof(int4) * ch.size(), hipMemcpyHostToDevice, g->stream));
    GG_HIP(hipStreamSynchronize(g->stream));
    g->ms_chunk_count = (int64_t)ch.size();
  }
  *out = g->ms_chunks.as<int4>();
  *count = g->ms_chunk_count;
  return GG_OK;
}

// next -> seen / F_next / union queue / depths.  perm: internal -> caller id
// (NULL: same).  Clears next.
__global__ void __launch_bounds__(256)
k_ms_update(const int32_t *__restrict__ ro, unsigned long long *__restrict__ next, unsigned long long *__restrict__ seen,
            unsigned long long *__restrict__ Fn, int32_t n, int32_t k, const int32_t *__restrict__ perm,
            int32_t level1, int32_t *__restrict__ depth, int32_t *__restrict__ queue, int32_t *__restrict__ qsize,
            unsigned long long *__restrict__ counters) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  unsigned long long mf = 0;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
    const int64_t vi = base + threadIdx.x;
    const int32_t v = (int32_t)vi;
    unsigned long long nx = 0;
    if (vi < n) {
      nx = next[v];
      Fn[v] = nx;
      if (nx) {
        next[v] = 0ull;
        seen[v] |= nx;
        mf += (unsigned long long)(ro[v + 1] - ro[v]);
      }
    }
    block_append<256>(nx != 0, v, queue, qsize);
    unsigned long long any = wave_or(nx);
    const int64_t pv = vi < n ? (perm ? perm[v] : v) : 0;
    while (any) {
      const int i = __builtin_ctzll(any);
      any &= any - 1;
      if ((nx >> i) & 1ull) depth[(int64_t)i * n + pv] = level1;
    }
  }
  block_counter_add<256>(&counters[kMsMfSlot], mf);
}

// Sources: seen, F and depth 0 for each instance; the union queue of the
// distinct sources (in order of first appearance) and its out-degree sum.
// One wave, lane i for instance i (k <= 64): a source repeated by several
// instances joins the queue from its first lane, and its bits are combined
// over those lanes before one store per word.
__global__ void __launch_bounds__(64)
k_ms_sources(const int32_t *__restrict__ ro, const int32_t *__restrict__ src, int32_t k,
             unsigned long long *__restrict__ seen, unsigned long long *__restrict__ F, int32_t n,
             const int32_t *__restrict__ perm, int32_t *__restrict__ depth, int32_t *__restrict__ queue,
             int32_t *__restrict__ qsize, unsigned long long *__restrict__ counters) {
  const int i = threadIdx.x;
  const bool on = i < k;
  const int32_t s = on ? src[i] : -1;
  // The lanes holding the same source as this lane, and whether this lane is the first of them.
  unsigned long long same = 0;
  for (int j = 0; j < k; j++) {
    const int32_t sj = __shfl(s, j);
    if (sj == s) same |= 1ull << j;
  }
  const bool first = on && (same & ((1ull << i) - 1ull)) == 0;
  const unsigned long long lead = ballot(first);
  const int32_t pos = __popcll(lead & ((1ull << i) - 1ull));
  unsigned long long deg = 0;
  if (first) {
    queue[pos] = s;
    seen[s] = same;  // seen and F start at zero
    F[s] = same;
    deg = (unsigned long long)(ro[s + 1] - ro[s]);
  }
  if (on) depth[(int64_t)i * n + (perm ? perm[s] : s)] = 0;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) deg += __shfl_xor(deg, off);
  if (i == 0) {
    qsize[0] = __popcll(lead);
    qsize[1] = 0;
    counters[kMsMfSlot] = deg;
  }
}

// One pass of up to 64 instances; depth: instance-major k x n on the device.
int ms_bfs_pass(gg_graph *g, const int32_t *src_int, int32_t k, int32_t *depth, gg_stats *acc) {
  hipStream_t s = g->stream;
  const int32_t n = g->n;
  const int64_t m = g->m;
  LbsBuffers lb;
  GG_TRY(lbs_buffers(g, &lb));
  GG_TRY(g->msbfs[0].ensure(8 * (size_t)n));  // seen
  GG_TRY(g->msbfs[1].ensure(16 * (size_t)n));  // F[2]
  GG_TRY(g->msbfs[2].ensure(8 * (size_t)n));  // next
  GG_TRY(g->msbfs[3].ensure(4 * (size_t)(k + 4)));  // sources, queue sizes
  const int4 *chunks = nullptr;
  int64_t nchunks = 0;
  GG_TRY(ms_long_chunks(g, &chunks, &nchunks));
  unsigned long long *seen = g->msbfs[0].as<unsigned long long>();
  unsigned long long *F[2] = {g->msbfs[1].as<unsigned long long>(), g->msbfs[1].as<unsigned long long>() + n};
  unsigned long long *next = g->msbfs[2].as<unsigned long long>();
  int32_t *qsz = g->msbfs[3].as<int32_t>();  // [0], [1] queue sizes, then the sources
  int32_t *d_src = qsz + 4;
  int32_t *queue[2] = {g->frontier[0].as<int32_t>(), g->frontier[1].as<int32_t>()};
  unsigned long long *ctr = g->counters.as<unsigned long long>();
  const int32_t *perm = g->reordered ? g->re_order.as<int32_t>() : nullptr;
  const unsigned long long full = k == 64 ? ~0ull : ((1ull << k) - 1ull);

  GG_HIP(hipMemcpyAsync(d_src, src_int, 4 * (size_t)k, hipMemcpyHostToDevice, s));
  GG_HIP(fill_async<unsigned long long>(seen, 0ull, n, s));
  GG_HIP(fill_async<unsigned long long>(F[0], 0ull, 2 * (int64_t)n, s));
  GG_HIP(fill_async<unsigned long long>(next, 0ull, n, s));
  GG_HIP(fill_async<int32_t>(depth, -1, (int64_t)k * n, s));
  GG_HIP(fill_async<unsigned long long>(ctr, 0ull, kCounterWords, s));
  hipLaunchKernelGGL(k_ms_sources, dim3(1), dim3(64), 0, s, g->dg.row_offsets, d_src, k, seen, F[0], n, perm, depth,
                     queue[0], qsz, ctr);
  GG_HIP(hipGetLastError());
  GG_TRY(counters_to_host_async(g, kCounterSlots, 0, qsz, 8));
  GG_HIP(hipStreamSynchronize(s));
  int64_t nf = reinterpret_cast<const int32_t *>(g->h64() + kCounterSlots)[0];
  int64_t mf = g->h64()[kMsMfSlot];
  int32_t level = 0;
  int64_t edges = 0, pulls = 0;
  MsPushOp op{};
  op.ro = g->dg.row_offsets;
  op.col = g->dg.column_indices;
  op.seen = seen;
  op.next = next;
  LbsPhaseTimer lbs_timer(g->prof, "bfs_batch", "push");
  while (nf > 0) {
    const int cur = level & 1;
    const bool bottom_up = mf > m / 20;
    // Clear the counters this level adds to and the next queue's size.
    GG_HIP(fill_async<int32_t>(qsz + (cur ^ 1), 0, 1, s));
    GG_HIP(slot_resets_async(ctr, {{kMsMfSlot, 0ull}, {kMsExaminedSlot, 0ull}}, s));
    if (bottom_up) {
      TimedRegion t(g->prof, "bfs_batch.bottom_up");
      hipLaunchKernelGGL(k_ms_bottom_up, dim3(grid_for(((int64_t)n + 63) / 64 * 64, 256, 8192)), dim3(256), 0, s,
                         g->dg.column_offsets, g->dg.row_indices, F[cur], seen, next, n, full, chunks, nchunks, ctr);
      pulls++;
    } else {
      op.F = F[cur];
      lbs_launch(op, false, queue[cur], qsz + cur, nf, (int64_t)n + m, lb, queue[cur ^ 1], qsz + (cur ^ 1), s,
                 lbs_timer);
      edges += mf;
    }
    {
      // At most 1024 workgroups: each appends to the union queue with one
      // same-address atomic (4096: 2.57-2.60 ms per 128 sources, 1024:
      // 2.49-2.53, 256: 2.93-3.00).
      TimedRegion t(g->prof, "bfs_batch.update");
      hipLaunchKernelGGL(k_ms_update, dim3(grid_for(n, 256, 1024)), dim3(256), 0, s, g->dg.row_offsets, next, seen,
                         F[cur ^ 1], n, k, perm, level + 1, depth, queue[cur ^ 1], qsz + (cur ^ 1), ctr);
    }
    GG_HIP(hipGetLastError());
    GG_TRY(counters_to_host_async(g, kCounterSlots, 0, qsz, 8));
    GG_HIP(hipStreamSynchronize(s));
    nf = reinterpret_cast<const int32_t *>(g->h64() + kCounterSlots)[cur ^ 1];
    mf = g->h64()[kMsMfSlot];
    if (bottom_up) edges += g->h64()[kMsExaminedSlot];
    level++;
  }
  acc->iterations = std::max(acc->iterations, level);
  acc->pull_iterations = std::max<int32_t>(acc->pull_iterations, (int32_t)pulls);
  acc->edges_inspected += edges;
  return GG_OK;
}

}  // namespace

int batch_bfs(gg_graph *g, const int32_t *sources, int32_t num, int32_t *depth_out, uint32_t flags, gg_stats *acc) {
  const size_t n = (size_t)g->n;
  std::vector<int32_t> src_int(sources, sources + num);
  for (auto &v : src_int) v = map_vertex(g, v);
  for (int32_t done = 0; done < num; done += kBfsBatchMax) {
    const int32_t k = std::min(kBfsBatchMax, num - done);
    int32_t *d = nullptr;
    if (flags & GG_OUT_DEVICE) {
      d = depth_out + (size_t)done * n;
    } else {
      GG_TRY(g->vtx[2].ensure(4 * n * (size_t)k));
      d = g->vtx[2].as<int32_t>();
    }
    ScopedEvents ev;
    GG_HIP(ev.create());
    GG_HIP(hipEventRecord(ev.e0, g->stream));
    GG_TRY(ms_bfs_pass(g, src_int.data() + done, k, d, acc));
    GG_HIP(hipEventRecord(ev.e1, g->stream));
    if (!(flags & GG_OUT_DEVICE))
      GG_HIP(hipMemcpyAsync(depth_out + (size_t)done * n, d, 4 * n * (size_t)k, hipMemcpyDeviceToHost, g->stream));
    GG_HIP(hipStreamSynchronize(g->stream));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ev.e0, ev.e1);
    acc->gpu_ms += ms;
  }
  return GG_OK;
}

}  // namespace gg
