// Direction-optimising BFS for gfx950 (default path of gg_bfs).
//
// Same result contract as the reference's BFS program (Algorithms/BFS/bfs.h,
// checked against CPUBFS bfs.cu:57-111): depth[src] = 0, unreached = -1.
//
//   top-down  (queue -> queue): LBS expand over the frontier's out-edges
//             (GASengine::lbs_launch with BfsPushOp); first touch wins an
//             atomic test-and-set in the visited bitmap and stamps the depth.
//   bottom-up (bitmap -> bitmap): a lane per unvisited vertex looks for a
//             parent in the frontier bitmap among its in-edges, stopping at
//             the first hit; the first four in-neighbours of every vertex are
//             read from a dense 16-byte-per-vertex copy, the hubs' frontier
//             bits from LDS (k_bfs_bottom_up).  A wave owns its 64 vertices'
//             bitmap words and writes them with plain stores.
//   switch    top-down -> bottom-up when m_f > m_u / alpha and the frontier
//             grows; back when it shrinks below n / beta (Beamer et al.).
//
// Byte model (per level, F frontier, E edges examined, N newly visited,
// U unvisited vertices scanned): top-down 12F + 8E + 8N; bottom-up
// 8U + 4E + n/8 + 8N.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>

#include "internal.h"

namespace gg {
namespace {

using namespace GASengine;
constexpr int kLoopLevels = 6;  // levels per replay of the captured loop (even)
// A top-down level over a small frontier (at most kTierItems vertices and
// kTierEdges out-edges) is one tiered launch instead of the four LBS launches.
constexpr int32_t kTierItems = 2048;
constexpr int64_t kTierEdges = 65536;

// counters: [0] next frontier size, [1] out-degree sum of the new frontier
// (m_f), [2] edges examined, [3] unvisited vertices scanned.
struct BfsPushOp {
  static constexpr bool kItemValues = false;
  const int32_t *ro;
  const int32_t *col;
  int32_t *depth;
  uint32_t *visited;
  unsigned long long *counters;
  int32_t next_depth;
  const int32_t *level;  // device-driven loop: depth = *level + 1 (else next_depth)
  unsigned long long mf;

  __device__ __forceinline__ void prologue(int32_t) {}
  __device__ __forceinline__ int32_t degree(int32_t v) { return ro[v + 1] - ro[v]; }
  __device__ __forceinline__ int32_t row_begin(int32_t v) { return ro[v]; }
  __device__ __forceinline__ int32_t item_value(int32_t) { return 0; }
  // Staged form for k_lbs_expand: all of a thread's loads for its edges of a
  // tile go out before any test-and-set: the column ids, then each target's
  // visited word and row bounds (the bounds are used only if this edge
  // discovers the target).  Bits are only set during the level, so an early
  // read of the word can only miss a set bit, and then the atomic decides.
  struct Loads {
    int32_t nb;
    uint32_t word;
    int32_t r0, r1;
  };
  __device__ __forceinline__ Loads load(int32_t, int32_t, int32_t slot) {
    const int32_t nb = col[slot];
    return {nb, visited[nb >> 5], ro[nb], ro[nb + 1]};
  }
  __device__ __forceinline__ bool apply(int32_t, const Loads &l, int32_t &out) {
    const int32_t nb = l.nb;
    const uint32_t m = 1u << (nb & 31);
    if (l.word & m) return false;
    if (atomicOr(&visited[nb >> 5], m) & m) return false;
    depth[nb] = level ? *level + 1 : next_depth;
    mf += (unsigned long long)(l.r1 - l.r0);
    out = nb;
    return true;
  }
  __device__ __forceinline__ bool edge(int32_t v, int32_t val, int32_t slot, int32_t &out) {
    return apply(v, load(v, val, slot), out);
  }
  __device__ __forceinline__ void finish() { block_counter_add<kExpandBlock>(&counters[1], mf); }
};

// Bottom-up level: each unvisited vertex looks for a parent (a frontier
// vertex among its in-neighbours), stopping at the first hit.  One lane per
// vertex, a wave per group of 64 consecutive vertices.  The first four
// in-neighbours of every vertex also sit in a dense array of 16 bytes per
// vertex (k_first_in_edges, built once per graph): a wave reads those of its
// 64 vertices with one coalesced load instead of touching up to 64 rows, and
// with the in-rows sorted by source id and the vertices renumbered by
// degree they are the likeliest parents, so most vertices need nothing else.
// The rest continue through their rows, 4 in-edges per step, or with the
// whole wave when the row is longer than kBuSerialMax.  The frontier bits of
// the kBuHotBits lowest-numbered vertices (the hubs) are copied to LDS at
// the start, so most probes are LDS reads.  Software-pipelined over a wave's
// groups: while group p is probed, the first in-neighbours of group p + 1 and
// the visited word and in-row offsets of group p + 2 are in flight (only this
// wave writes those groups' visited words).  The wave owns its groups'
// bitmap words and writes them with plain stores; vertices without in-edges
// were marked visited at the start of the run.
//
// ro: CSR offsets for the out-degree sum of the new frontier; nullptr on an
// undirected (symmetric) graph, where the in-row length is the out-degree.
constexpr int kBuBlock = 256;
constexpr int kBuHotWords = 4096;  // 16 KB of LDS
constexpr int32_t kBuHotBits = kBuHotWords * 32;
constexpr int kBuGridMax = 2048;
constexpr int32_t kBuSerialMax = 512;
static_assert(kBuHotWords % kBuBlock == 0, "the LDS copy takes the same number of words per thread");

__global__ void __launch_bounds__(kBuBlock)
k_bfs_bottom_up(const int32_t *__restrict__ co, const int32_t *__restrict__ ri, const int4 *__restrict__ first4,
                const int32_t *__restrict__ ro, int32_t *__restrict__ depth, uint32_t *__restrict__ visited,
                const uint32_t *__restrict__ front, uint32_t *__restrict__ next, int32_t n, int32_t next_depth,
                unsigned long long *__restrict__ counters, const int32_t *__restrict__ skip,
                const int32_t *__restrict__ level) {
  if (skip && *skip) return;
  if (level) next_depth = *level + 1;
  const int32_t words = (n + 31) >> 5;
  __shared__ uint32_t s_hot[kBuHotWords];
  {
    // All of a thread's loads in flight before its LDS stores.
    constexpr int kPer = kBuHotWords / kBuBlock;
    uint32_t t[kPer];
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      const int i = (int)threadIdx.x + k * kBuBlock;
      t[k] = i < words ? front[i] : 0u;
    }
#pragma unroll
    for (int k = 0; k < kPer; k++) s_hot[threadIdx.x + k * kBuBlock] = t[k];
  }
  __syncthreads();
  auto fbit = [&](int32_t u) -> uint32_t {
    return ((u < kBuHotBits ? s_hot[u >> 5] : front[u >> 5]) >> (u & 31)) & 1u;
  };
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const int lane = lane_id();
  unsigned long long found_cnt = 0, mf = 0, examined = 0, scanned = 0;
  int64_t p = (((int64_t)blockIdx.x * blockDim.x) + threadIdx.x) >> 6;
  // In-row offsets of this lane's vertex of group q; b = e = 0 when it is
  // visited or beyond n (nothing to scan).
  auto rows = [&](int64_t q, int32_t &rb, int32_t &re) {
    rb = re = 0;
    const int64_t v = q * 64 + lane;
    if (v < n) {
      const uint32_t w = visited[2 * q + (lane >> 5)];
      if (!((w >> (lane & 31)) & 1u)) {
        rb = co[v];
        re = co[v + 1];
      }
    }
  };
  auto firsts = [&](int64_t q, int32_t rb, int32_t re) {
    int4 f = make_int4(-1, -1, -1, -1);
    if (re > rb) f = first4[q * 64 + lane];
    return f;
  };
  int32_t b0, e0, b1, e1;
  rows(p, b0, e0);
  int4 f0 = firsts(p, b0, e0);
  rows(p + nw, b1, e1);
  for (; p * 64 < n; p += nw) {
    const int4 f1 = firsts(p + nw, b1, e1);
    int32_t b2, e2;
    rows(p + 2 * nw, b2, e2);
    const int32_t v = (int32_t)(p * 64 + lane);
    const bool todo = e0 > b0;
    if (todo) scanned++;
    bool found = false;
    if (todo) {
      // f0.x is an in-neighbour (row not empty); the others are -1 past the row's end
      found = (fbit(f0.x) | (f0.y >= 0 ? fbit(f0.y) : 0u) | (f0.z >= 0 ? fbit(f0.z) : 0u) |
               (f0.w >= 0 ? fbit(f0.w) : 0u)) != 0;
      examined += (unsigned long long)min(4, e0 - b0);
    }
    const bool more = todo && !found && e0 - b0 > 4;
    const bool coop = more && (e0 - b0) > kBuSerialMax;
    if (more && !coop) {
      int32_t k = b0 + 4;
      for (; k + 4 <= e0 && !found; k += 4) {
        const int32_t u0 = ri[k], u1 = ri[k + 1], u2 = ri[k + 2], u3 = ri[k + 3];
        found = (fbit(u0) | fbit(u1) | fbit(u2) | fbit(u3)) != 0;
        examined += 4;
      }
      for (; k < e0 && !found; k++) {
        found = fbit(ri[k]) != 0;
        examined++;
      }
    }
    uint64_t cm = ballot(coop);
    while (cm) {
      const int l = __builtin_ctzll(cm);
      cm &= cm - 1;
      const int32_t wb = __shfl(b0, l) + 4, we = __shfl(e0, l);
      bool hit = false;
      for (int32_t k0 = wb; k0 < we; k0 += kWave) {
        const int32_t k = k0 + lane;
        const bool h = k < we && fbit(ri[k]);
        if (lane == 0) examined += (unsigned long long)min(kWave, we - k0);
        if (ballot(h)) {
          hit = true;
          break;
        }
      }
      if (lane == l) found = hit;
    }
    const uint64_t mask = ballot(found);
    if (found) {
      depth[v] = next_depth;
      mf += (unsigned long long)(ro ? ro[v + 1] - ro[v] : e0 - b0);
    }
    if (lane < 2) {
      const int32_t ww = (int32_t)(2 * p + lane);
      const uint32_t bits = lane ? (uint32_t)(mask >> 32) : (uint32_t)mask;
      if (ww < words) {
        next[ww] = bits;
        if (bits) visited[ww] |= bits;
      }
    }
    if (lane == 0) found_cnt += (unsigned long long)__popcll(mask);
    b0 = b1;
    e0 = e1;
    f0 = f1;
    b1 = b2;
    e1 = e2;
  }
  block_counter_add<kBuBlock>(&counters[0], found_cnt);
  block_counter_add<kBuBlock>(&counters[1], mf);
  block_counter_add<kBuBlock>(&counters[2], examined);
  block_counter_add<kBuBlock>(&counters[3], scanned);
}

inline int bottom_up_grid(int32_t n) { return grid_for(((int64_t)n + 63) / 64, kBuBlock / kWave, kBuGridMax); }

// first4[v] = the first four in-neighbours of v (CSC row order), -1 past the
// end of a shorter row.
__global__ void __launch_bounds__(256)
k_first_in_edges(const int32_t *__restrict__ co, const int32_t *__restrict__ ri, int32_t n, int4 *__restrict__ first4) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
    const int32_t b = co[v], d = co[v + 1] - b;
    first4[v] = make_int4(d > 0 ? ri[b] : -1, d > 1 ? ri[b + 1] : -1, d > 2 ? ri[b + 2] : -1, d > 3 ? ri[b + 3] : -1);
  }
}

// Bit v of the result: vertex v has no in-edges (CSC row empty).  Bits at
// and beyond n stay clear.
__global__ void __launch_bounds__(256)
k_no_in_edges(const int32_t *__restrict__ co, int32_t n, uint32_t *__restrict__ bits) {
  const int32_t words = (n + 31) >> 5;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const int lane = lane_id();
  for (int64_t p = (((int64_t)blockIdx.x * blockDim.x) + threadIdx.x) >> 6; p * 64 < n; p += nwaves) {
    const int64_t v = p * 64 + lane;
    const uint64_t mask = ballot(v < n && co[v + 1] == co[v]);
    const int64_t w = 2 * p + lane;
    if (lane < 2 && w < words) bits[w] = lane ? (uint32_t)(mask >> 32) : (uint32_t)mask;
  }
}

__global__ void k_queue_to_bits(const int32_t *__restrict__ queue, const int32_t *__restrict__ qsize,
                                uint32_t *__restrict__ bits, const int32_t *__restrict__ enable) {
  if (enable && !*enable) return;
  const int32_t nq = *qsize;
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += gridDim.x * blockDim.x) {
    const int32_t v = queue[i];
    atomicOr(&bits[v >> 5], 1u << (v & 31));
  }
}

__global__ void __launch_bounds__(256)
k_bits_to_queue(const uint32_t *__restrict__ bits, int32_t words, int32_t *__restrict__ queue,
                int32_t *__restrict__ qsize, const int32_t *__restrict__ enable) {
  if (enable && !*enable) return;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < words; base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t w = base + threadIdx.x;
    uint32_t word = w < words ? bits[w] : 0u;
    const int32_t cnt = __popc(word);
    const int32_t incl = wave_inclusive_scan(cnt);
    const int32_t total = __shfl(incl, kWave - 1);
    int32_t start = 0;
    if (lane_id() == 0 && total) start = atomicAdd(qsize, total);
    start = __shfl(start, 0) + incl - cnt;
    while (word) {
      const int bb = __builtin_ctz(word);
      word &= word - 1;
      queue[start++] = (int32_t)(w * 32 + bb);
    }
  }
}

// Conversions between the two frontier forms, one launch for both (the
// controller enables at most one).  Queue -> bitmap is rebuilt from the
// depths (vertices at depth == level form the current frontier): whole
// words written by a ballot, so no clearing pass and no atomics.
__global__ void __launch_bounds__(256)
k_bfs_convert(const int32_t *__restrict__ depth, const int32_t *__restrict__ level, const int32_t *__restrict__ q2b,
              const int32_t *__restrict__ b2q, int32_t n, uint32_t *__restrict__ bits, int32_t *__restrict__ queue,
              int32_t *__restrict__ qsize) {
  const int32_t words = (n + 31) >> 5;
  if (*q2b) {
    const int32_t lv = *level;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int lane = lane_id();
    for (int64_t p = (((int64_t)blockIdx.x * blockDim.x) + threadIdx.x) >> 6; p * 64 < n; p += nwaves) {
      const int64_t v = p * 64 + lane;
      const uint64_t mask = ballot(v < n && depth[v] == lv);
      const int64_t w = 2 * p + lane;
      if (lane < 2 && w < words) bits[w] = lane ? (uint32_t)(mask >> 32) : (uint32_t)mask;
    }
  } else if (*b2q) {
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < words; base += (int64_t)gridDim.x * blockDim.x) {
      const int64_t w = base + threadIdx.x;
      uint32_t word = w < words ? bits[w] : 0u;
      const int32_t cnt = __popc(word);
      const int32_t incl = wave_inclusive_scan(cnt);
      const int32_t total = __shfl(incl, kWave - 1);
      int32_t start = 0;
      if (lane_id() == 0 && total) start = atomicAdd(qsize, total);
      start = __shfl(start, 0) + incl - cnt;
      while (word) {
        const int bb = __builtin_ctz(word);
        word &= word - 1;
        queue[start++] = (int32_t)(w * 32 + bb);
      }
    }
  }
}

// Graphs of at most kSmallN vertices: the whole BFS in one workgroup and one
// launch (level loop inside the kernel, frontiers, their degree prefix and
// the visited bitmap in LDS), so a small graph costs one launch and one
// synchronisation instead of several launches per level.  Each level's
// frontier edges are spread over the 1024 threads by a binary search in the
// LDS degree prefix (balanced whatever the degrees).  Results: depth (global),
// counters slot 0 = levels, 1 = reached vertices, 2 = out-degree sum of the
// reached vertices (= edges examined), every other stripe of those slots 0.
constexpr int kSmallN = 8192;
constexpr int kSmallBlock = 1024;
constexpr int kSmallPer = kSmallN / kSmallBlock;

__global__ void __launch_bounds__(kSmallBlock)
k_bfs_small(const int32_t *__restrict__ ro, const int32_t *__restrict__ col, int32_t n, int32_t src,
            int32_t *__restrict__ depth, unsigned long long *__restrict__ ctr) {
  __shared__ int32_t q[2][kSmallN];
  __shared__ int32_t off[kSmallN];
  __shared__ uint32_t vis[kSmallN / 32];
  __shared__ int32_t qn[2];
  __shared__ int32_t scan[kSmallBlock / kWave + 1];
  const int tid = threadIdx.x;
  for (int v = tid; v < n; v += kSmallBlock) depth[v] = v == src ? 0 : -1;
  for (int w = tid; w < kSmallN / 32; w += kSmallBlock) vis[w] = w == (src >> 5) ? 1u << (src & 31) : 0u;
  if (tid == 0) {
    q[0][0] = src;
    qn[0] = 1;
    qn[1] = 0;
  }
  __syncthreads();
  long long reached = 1, edges = 0;
  int32_t level = 0;
  for (;;) {
    const int cur = level & 1;
    const int32_t nq = qn[cur];
    if (nq == 0) break;
    // Degree prefix of the frontier: kSmallPer consecutive items per thread.
    const int32_t i0 = tid * kSmallPer;
    int32_t deg[kSmallPer];
    int32_t sum = 0;
#pragma unroll
    for (int j = 0; j < kSmallPer; j++) {
      deg[j] = 0;
      if (i0 + j < nq) {
        const int32_t u = q[cur][i0 + j];
        deg[j] = ro[u + 1] - ro[u];
      }
      sum += deg[j];
    }
    int32_t total;
    int32_t ex = block_exclusive_scan<kSmallBlock>(sum, scan, &total);
#pragma unroll
    for (int j = 0; j < kSmallPer; j++) {
      if (i0 + j < nq) off[i0 + j] = ex;
      ex += deg[j];
    }
    if (tid == 0) qn[cur ^ 1] = 0;
    __syncthreads();
    edges += total;
    for (int32_t k = tid; k < total; k += kSmallBlock) {
      int32_t lo = 0, hi = nq - 1;  // last item with off <= k
      while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (off[mid] <= k)
          lo = mid;
        else
          hi = mid - 1;
      }
      const int32_t u = q[cur][lo];
      const int32_t v = col[ro[u] + (k - off[lo])];
      const uint32_t bit = 1u << (v & 31);
      if (!(vis[v >> 5] & bit)) {
        const uint32_t old = atomicOr(&vis[v >> 5], bit);
        if (!(old & bit)) {
          depth[v] = level + 1;
          q[cur ^ 1][atomicAdd(&qn[cur ^ 1], 1)] = v;
        }
      }
    }
    __syncthreads();
    reached += qn[cur ^ 1];
    level++;
  }
  if (tid < kStripes * 3) {
    const int q0 = tid / 3, slot = tid % 3;
    unsigned long long v = 0;
    if (q0 == 0) v = slot == 0 ? (unsigned long long)level : slot == 1 ? (unsigned long long)reached
                                                                      : (unsigned long long)edges;
    ctr[q0 * kCounterSlots + slot] = v;
  }
}

// Kinds of a level (what its slot in a scheduled replay launches); bit k of
// a slot's mask allows kind k.
enum BfsKind : int {
  kBfsBottomUp = 1,      // bottom-up
  kBfsBottomUpConv = 2,  // queue -> bitmap conversion, bottom-up
  kBfsTier = 3,          // small top-down (tiered launch)
  kBfsTierConv = 4,      // bitmap -> queue conversion, small top-down
  kBfsLbs = 5,           // top-down by LBS
  kBfsLbsConv = 6,       // bitmap -> queue conversion, top-down by LBS
};
constexpr int kBfsMaskAll = 0x7e;

// State of the device-driven loop.  One controller launch per level closes
// the previous level (folds its counters, the host loop's bookkeeping),
// clears the counters for the new one and takes its direction decision; a
// last launch at the end of the captured sequence only closes.
struct BfsCtl {
  int32_t level;      // levels done; the next level writes depth level + 1
  int32_t bottom_up;  // direction of the level in flight
  int32_t q2b, b2q;   // conversions before this level
  int32_t bu_skip;    // 1 unless this level runs bottom-up
  int32_t push_size;  // frontier items of a top-down level by LBS, else 0
  int32_t tier_size;  // frontier items of a small top-down level (tiered launch), else 0
  int32_t done, pull_levels, pending;
  int32_t stall;      // a level's kind was not in its slot of a scheduled replay: it has not started
  long long nf, prev_nf, mf, mu, reached, reached_edges, examined;
  double bytes;
  uint8_t kinds[kSpecMax];  // kind of each level started (BfsKind), the first kSpecMax levels
};
static_assert(sizeof(BfsCtl) % 8 == 0, "read back in 8-byte words");

// The start of a run in one launch, every element at its final initial value
// (no fill then patch, so no ordering between threads is needed): depth = -1
// except depth[src] = 0; visited = the vertices without in-edges (no level
// can reach them) plus src; the counters zero
// except [1] = out-degree of the source and the first queue size (low half
// of slot 8) = 1; queue[0] = src; and, for the device-driven loop, its state
// block.
__global__ void k_bfs_start(int32_t *depth, int32_t n, uint32_t *visited, const uint32_t *noin, int32_t words,
                            unsigned long long *ctr, int32_t *queue, const int32_t *ro, int32_t src, BfsCtl *c,
                            int64_t m, double bytes0) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const long long mf = (long long)(ro[src + 1] - ro[src]);
  fill_except_one(depth, n, -1, src, 0, t0, stride);
  for (int64_t i = t0; i < words; i += stride) visited[i] = noin[i] | (i == (src >> 5) ? 1u << (src & 31) : 0u);
  for (int64_t i = t0; i < kCounterWords; i += stride)
    ctr[i] = i == 1 ? (unsigned long long)mf : i == 8 ? 1ull : 0ull;  // slot 8: int32 queue sizes {1, 0}
  if (t0 == 0) {
    queue[0] = src;
    if (c) {
      c->level = 0;
      c->bottom_up = 0;
      c->q2b = c->b2q = 0;
      c->bu_skip = 1;
      c->push_size = 0;
      c->tier_size = 0;
      c->done = 0;
      c->pull_levels = 0;
      c->pending = 0;
      c->stall = 0;
      c->nf = 1;
      c->prev_nf = 0;
      c->mf = mf;
      c->mu = m - mf;
      c->reached = 1;
      c->reached_edges = mf;
      c->examined = 0;
      c->bytes = bytes0;
    }
  }
}

// qsize_cur: the queue the level in flight wrote (its next queue) and the
// new level reads; qsize_nxt: the new level's output queue.  mask: the level
// kinds this slot launches kernels for (kBfsMaskAll in the general loop).  A
// level whose kind is not in the mask is not started: the controller sets
// `stall`, the rest of the replay does nothing, and the general loop's first
// controller (which clears `stall`) takes the same decision again.
__global__ void __launch_bounds__(64)
k_bfs_ctrl(BfsCtl *c, unsigned long long *__restrict__ ctr, int32_t *__restrict__ qsize_cur,
           int32_t *__restrict__ qsize_nxt, int32_t n, int32_t alpha, int32_t beta, int32_t max_levels,
           int close_only, int mask) {
  __shared__ int s_go;
  if (threadIdx.x == 0) {
    const bool stalled = c->stall && mask != kBfsMaskAll;
    if (!stalled) {
      c->stall = 0;
      if (c->pending && !c->done) {  // close the level in flight
        const long long E = (long long)stripe_sum(ctr + 2);
        long long N;
        if (c->bottom_up) {
          N = (long long)stripe_sum(ctr + 0);
          c->bytes += 8.0 * (double)stripe_sum(ctr + 3) + 4.0 * E + n / 8.0 + 8.0 * N;
          c->pull_levels++;
        } else {
          N = *qsize_cur;
          c->bytes += 12.0 * c->nf + 8.0 * E + 8.0 * N;
        }
        c->examined += E;
        c->prev_nf = c->nf;
        c->nf = N;
        c->mf = (long long)stripe_sum(ctr + 1);
        c->mu -= c->mf;
        c->reached += N;
        c->reached_edges += c->mf;
        c->level++;
        if (N == 0 || c->level >= max_levels) c->done = 1;
      }
      c->pending = 0;
      c->q2b = c->b2q = 0;
      c->bu_skip = 1;
      c->push_size = 0;
      c->tier_size = 0;
    }
    s_go = !stalled && !close_only && !c->done;
  }
  __syncthreads();
  if (!s_go) return;
  // counters of the new level: [0..3] in every stripe, and its output queue
  {
    const int q = threadIdx.x;  // 64 threads: 16 stripes x 4 counters
    ctr[(q >> 2) * kCounterSlots + (q & 3)] = 0ull;
  }
  if (threadIdx.x == 0) {
    *qsize_nxt = 0;
    int bu = c->bottom_up, q2b = 0, b2q = 0;
    if (!bu && c->mf > c->mu / alpha && c->nf > c->prev_nf) {
      q2b = 1;
      bu = 1;
    } else if (bu && c->nf < n / beta && c->nf < c->prev_nf) {
      b2q = 1;
      bu = 0;
    }
    const bool tier = !bu && c->nf <= kTierItems && c->mf <= kTierEdges;
    const int kind = bu ? (q2b ? kBfsBottomUpConv : kBfsBottomUp)
                        : tier ? (b2q ? kBfsTierConv : kBfsTier) : (b2q ? kBfsLbsConv : kBfsLbs);
    if (!((mask >> kind) & 1)) {
      c->stall = 1;  // the flags stay cleared: every kernel of the replay returns
      return;
    }
    if (c->level < kSpecMax) c->kinds[c->level] = (uint8_t)kind;
    c->bottom_up = bu;
    c->q2b = q2b;
    c->b2q = b2q;
    if (b2q) *qsize_cur = 0;  // k_bfs_convert appends
    if (bu)
      c->bu_skip = 0;
    else if (tier)
      c->tier_size = (int32_t)c->nf;
    else
      c->push_size = (int32_t)c->nf;
    c->pending = 1;
  }
}

}  // namespace

int bfs_direction_optimizing(gg_graph *g, int32_t src, int32_t *depth_out, uint32_t flags, gg_stats *stats) {
  GG_TRY(activate(g));
  const int32_t n = g->n;
  const int32_t words = (n + 31) / 32;
  hipStream_t s = g->stream;
  LbsBuffers lb;
  GG_TRY(lbs_buffers(g, &lb));
  if (n <= kSmallN && g->opt.iter_num == INT_MAX) {
    int32_t *depth = depth_out;
    if (!(flags & GG_OUT_DEVICE)) {
      GG_TRY(g->vtx[0].ensure(4 * (size_t)n));
      depth = g->vtx[0].as<int32_t>();
    } else {
      GG_CHECK_ARG(depth_out, "depth is NULL");
    }
    g->prof.enabled = (flags & GG_INSTRUMENT) != 0;
    unsigned long long *ctr = g->counters.as<unsigned long long>();
    hipEvent_t e0, e1;
    GG_HIP(g->bfs_ev.get(&e0, &e1));
    GG_HIP(hipEventRecord(e0, s));
    {
      TimedRegion t(g->prof, "bfs.small");
      hipLaunchKernelGGL(k_bfs_small, dim3(1), dim3(kSmallBlock), 0, s, g->dg.row_offsets, g->dg.column_indices, n,
                         src, depth, ctr);
      GG_HIP(hipGetLastError());
    }
    GG_HIP(hipEventRecord(e1, s));
    GG_TRY(read_counters(g, 3));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    const int64_t E = g->h64()[2], N = g->h64()[1];
    g->prof.add_bytes("bfs.small", 12.0 * N + 8.0 * E + 8.0 * N + 4.0 * n);
    finish_profile(g);
    if (stats) {
      std::memset(stats, 0, sizeof(*stats));
      stats->iterations = (int32_t)g->h64()[0];
      stats->reached_vertices = N;
      stats->reached_edges = E;
      stats->edges_inspected = E;
      stats->alg_bytes = 12.0 * N + 8.0 * E + 8.0 * N + 4.0 * n;
      stats->gpu_ms = ms;
    }
    if (flags & GG_OUT_DEVICE) return GG_OK;
    return copy_out(g, depth_out, depth, 4 * (size_t)n, flags);
  }
  GG_TRY(g->bits[0].ensure(4 * (size_t)words));  // visited
  GG_TRY(g->bits[1].ensure(4 * (size_t)words));  // frontier bitmaps (ping-pong)
  GG_TRY(g->bits[2].ensure(4 * (size_t)words));
  int32_t *depth;
  if (flags & GG_OUT_DEVICE) {
    GG_CHECK_ARG(depth_out, "depth is NULL");
    depth = depth_out;
  } else {
    GG_TRY(g->vtx[0].ensure(4 * (size_t)n));
    depth = g->vtx[0].as<int32_t>();
  }
  uint32_t *visited = g->bits[0].as<uint32_t>();
  uint32_t *fbits[2] = {g->bits[1].as<uint32_t>(), g->bits[2].as<uint32_t>()};
  int32_t *queue[2] = {g->frontier[0].as<int32_t>(), g->frontier[1].as<int32_t>()};
  // counters (striped, GASengine::block_counter_add): [0..3] per level (see
  // BfsPushOp); slots 8, 9 hold the int32 queue sizes (stripe 0 only)
  unsigned long long *ctr = g->counters.as<unsigned long long>();
  int32_t *qsz[2] = {reinterpret_cast<int32_t *>(ctr + 8), reinterpret_cast<int32_t *>(ctr + 9)};
  g->prof.enabled = (flags & GG_INSTRUMENT) != 0;
  LbsPhaseTimer lbs_timer(g->prof, "bfs", "top_down");

  hipEvent_t e0, e1;
  GG_HIP(g->bfs_ev.get(&e0, &e1));
  GG_HIP(hipEventRecord(e0, s));
  BfsCtl *ctl = nullptr;  // state block of the device-driven loop
  if (!g->prof.enabled) {
    GG_TRY(g->loop_ctl.ensure(sizeof(BfsCtl)));
    ctl = g->loop_ctl.as<BfsCtl>();
  }
  // Vertices without in-edges, computed once per graph: every run starts
  // with them visited.
  if (!g->noin_ready) {
    GG_TRY(g->noin.ensure(4 * (size_t)words));
    hipLaunchKernelGGL(k_no_in_edges, dim3(grid_for((int64_t)n, 256, 2048)), dim3(256), 0, s, g->dg.column_offsets, n,
                       g->noin.as<uint32_t>());
    GG_HIP(hipGetLastError());
    g->noin_ready = true;
  }
  const uint32_t *noin = g->noin.as<uint32_t>();
  // The first four in-neighbours of every vertex, dense (bottom-up levels).
  if (!g->first4_ready) {
    GG_TRY(g->first4.ensure(16 * (size_t)n));
    hipLaunchKernelGGL(k_first_in_edges, dim3(grid_for((int64_t)n, 256, 8192)), dim3(256), 0, s,
                       g->dg.column_offsets, g->dg.row_indices, n, g->first4.as<int4>());
    GG_HIP(hipGetLastError());
    g->first4_ready = true;
  }
  const int4 *first4 = g->first4.as<int4>();
  {
    TimedRegion t(g->prof, "bfs.init");
    hipLaunchKernelGGL(k_bfs_start, dim3(grid_for(std::max<int64_t>(n, kCounterWords), 256, 8192)), dim3(256), 0, s,
                       depth, n, visited, noin, words, ctr, queue[0], g->dg.row_offsets, src, ctl, g->m,
                       4.0 * n + words * 8.0);
    GG_HIP(hipGetLastError());
  }
  g->prof.add_bytes("bfs.init", 4.0 * n + words * 8.0);
  int64_t mf = 0, mu = 0, nf = 1, prev_nf = 0;
  int64_t reached = 1, reached_edges = 0, examined_total = 0;
  double bytes = 4.0 * n + words * 8.0;
  const int alpha = g->opt.bfs_alpha, beta = g->opt.bfs_beta;
  bool bottom_up = false;
  int cur = 0;  // index of the current queue / frontier bitmap
  int levels = 0, pull_levels = 0;
  bool timed = false;  // the device-driven loop recorded e1 and synchronised past it
  if (!g->prof.enabled) {
    // Device-driven: replays of captured launch sequences.  The general loop
    // has kLoopLevels level slots (an even count, so a replay keeps the
    // parity of levels and ping-pong buffers; one capture per parity of its
    // first level).  Per slot: k_bfs_ctrl (closes the previous level, picks
    // the direction), k_bfs_convert, the bottom-up kernel, the LBS top-down
    // (over ctl->push_size) and the tiered top-down, each running or
    // returning on the controller's flags; a last k_bfs_ctrl closes the
    // replay's final level.  When the previous run on this graph took at
    // most kSpecMax levels, the first replay is that run's schedule: one slot
    // per level with only the kernels of the level's kind (a controller that
    // needs another kind stalls the replay, and the general loop goes on from
    // that level).  Per-kernel events (GG_INSTRUMENT) need the host loop
    // below.
    const std::vector<uint64_t> key = {
        (uint64_t)depth, (uint64_t)visited, (uint64_t)fbits[0], (uint64_t)fbits[1], (uint64_t)queue[0],
        (uint64_t)queue[1], (uint64_t)ctr, (uint64_t)ctl, (uint64_t)lb.scan_local, (uint64_t)lb.partials,
        (uint64_t)lb.info, (uint64_t)lb.parts, (uint64_t)lb.edge_counter, (uint64_t)lb.expand_grid, (uint64_t)lb.vt,
        (uint64_t)alpha, (uint64_t)beta, (uint64_t)g->opt.wave_degree, (uint64_t)g->opt.iter_num,
        (uint64_t)g->opt.block_degree};
    const int conv_grid = grid_for((int64_t)n, 256, 2048);  // strides by waves; idle at most levels
    // One level slot: slot parity c (its input queue / bitmap), the kinds
    // it launches for.
    auto launch_level = [&](int c, int mask) {
      const int x = c ^ 1;
      const auto has = [&](int k) { return (mask >> k) & 1; };
      hipLaunchKernelGGL(k_bfs_ctrl, dim3(1), dim3(64), 0, s, ctl, ctr, qsz[c], qsz[x], n, alpha, beta,
                         g->opt.iter_num, 0, mask);
      if (has(kBfsBottomUpConv) || has(kBfsTierConv) || has(kBfsLbsConv))
        hipLaunchKernelGGL(k_bfs_convert, dim3(conv_grid), dim3(256), 0, s, depth, (const int32_t *)&ctl->level,
                           (const int32_t *)&ctl->q2b, (const int32_t *)&ctl->b2q, n, fbits[c], queue[c], qsz[c]);
      if (has(kBfsBottomUp) || has(kBfsBottomUpConv))
        hipLaunchKernelGGL(k_bfs_bottom_up, dim3(bottom_up_grid(n)), dim3(kBuBlock), 0, s, g->dg.column_offsets,
                           g->dg.row_indices, first4, g->directed ? g->dg.row_offsets : nullptr, depth, visited,
                           fbits[c], fbits[x], n, 0, ctr, (const int32_t *)&ctl->bu_skip, (const int32_t *)&ctl->level);
      BfsPushOp op{g->dg.row_offsets, g->dg.column_indices, depth, visited, ctr, 0, &ctl->level, 0ull};
      if (has(kBfsLbs) || has(kBfsLbsConv))
        lbs_launch(op, false, queue[c], &ctl->push_size, n, (int64_t)n + g->m, lb, queue[x], qsz[x], s, [](int) {});
      if (has(kBfsTier) || has(kBfsTierConv))
        hipLaunchKernelGGL((k_tiered_expand<BfsPushOp, false>), dim3(grid_for(kTierItems, kExpandBlock)),
                           dim3(kExpandBlock), 0, s, op, queue[c], (const int32_t *)&ctl->tier_size, queue[x], qsz[x],
                           g->opt.wave_degree, g->opt.block_degree, lb.edge_counter);
    };
    // Captures `masks.size()` level slots starting at parity p, then the
    // closing controller (close_mask: kBfsMaskAll in the general loop, 0 in
    // a scheduled one, whose closing controller leaves a stall in place).
    auto capture = [&](CapturedLoop &L, int p, const std::vector<int> &masks, int close_mask) -> int {
      L.reset();
      hipGraph_t graph = nullptr;
      GG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      for (size_t lv = 0; lv < masks.size(); lv++) launch_level((int)((lv + p) & 1), masks[lv]);
      const int e = (int)((masks.size() + p) & 1);  // parity of the queue the last slot wrote
      hipLaunchKernelGGL(k_bfs_ctrl, dim3(1), dim3(64), 0, s, ctl, ctr, qsz[e], qsz[e ^ 1], n, alpha, beta,
                         g->opt.iter_num, 1, close_mask);
      const hipError_t le = hipGetLastError();
      const hipError_t ce = hipStreamEndCapture(s, &graph);
      hipError_t ie = hipErrorUnknown;
      if (le == hipSuccess && ce == hipSuccess) ie = hipGraphInstantiate(&L.exec, graph, nullptr, nullptr, 0);
      if (graph) (void)hipGraphDestroy(graph);
      if (ie != hipSuccess) {
        L.exec = nullptr;
        return set_error(GG_ERR_HIP, "BFS loop capture failed: %s / %s / %s", hipGetErrorString(le),
                         hipGetErrorString(ce), hipGetErrorString(ie));
      }
      L.key = key;
      return GG_OK;
    };
    auto general = [&](int p, hipGraphExec_t *out) -> int {
      CapturedLoop &L = g->bfs_loop[p];
      if (!L.exec || L.key != key) GG_TRY(capture(L, p, std::vector<int>(kLoopLevels, kBfsMaskAll), kBfsMaskAll));
      *out = L.exec;
      return GG_OK;
    };
    auto scheduled = [&](hipGraphExec_t *out) -> int {
      for (SpecLoop &sp : g->bfs_spec)
        if (sp.loop.exec && sp.loop.key == key && sp.sched == g->bfs_sched) {
          *out = sp.loop.exec;
          return GG_OK;
        }
      SpecLoop &sp = g->bfs_spec[g->bfs_spec_next];
      g->bfs_spec_next = (g->bfs_spec_next + 1) % kSpecCache;
      std::vector<int> masks;
      for (uint8_t k : g->bfs_sched) masks.push_back(1 << k);
      sp.sched.clear();
      GG_TRY(capture(sp.loop, 0, masks, 0));
      sp.sched = g->bfs_sched;
      *out = sp.loop.exec;
      return GG_OK;
    };
    bool first = true;
    int parity = 0;  // parity of the next level to run
    for (;;) {
      hipGraphExec_t exec = nullptr;
      if (first && !g->bfs_sched.empty())
        GG_TRY(scheduled(&exec));
      else
        GG_TRY(general(parity, &exec));
      first = false;
      GG_HIP(hipGraphLaunch(exec, s));
      GG_HIP(hipEventRecord(e1, s));  // the last replay's record times the run
      // The caller's hand-off, on the chance that this replay ends the run
      // (else it is enqueued again after the next one).
      if (flags & GG_OUT_DEVICE) GG_TRY(enqueue_handoff(g, depth));
      GG_TRY(counters_to_host_async(g, 0, 0, ctl, sizeof(BfsCtl)));
      GG_HIP(hipStreamSynchronize(s));
      BfsCtl h;
      std::memcpy(&h, g->h64() + kCounterSlots, sizeof(h));
      if (h.done) {
        if (flags & GG_OUT_DEVICE) g->handoff.done = g->handoff.armed;
        levels = h.level;
        pull_levels = h.pull_levels;
        reached = h.reached;
        reached_edges = h.reached_edges;
        examined_total = h.examined;
        bytes = h.bytes;
        nf = 0;
        timed = true;
        // Levels started: all that ran, the last of which found nothing.
        if (levels <= kSpecMax)
          g->bfs_sched.assign(h.kinds, h.kinds + levels);
        else
          g->bfs_sched.clear();
        break;
      }
      parity = h.level & 1;
    }
  } else {
    GG_TRY(read_counters(g, 16));
    mf = g->h64()[1];
    mu = g->m - mf;
    reached_edges = mf;
  }

  while (g->prof.enabled && nf > 0 && levels < g->opt.iter_num) {
    // Direction choice.
    if (!bottom_up && mf > mu / alpha && nf > prev_nf) {
      TimedRegion t(g->prof, "bfs.queue_to_bits");
      GG_HIP(fill_async<uint32_t>(fbits[cur], 0u, words, s));
      hipLaunchKernelGGL(k_queue_to_bits, dim3(grid_for(nf, 256, 4096)), dim3(256), 0, s, queue[cur], qsz[cur],
                         fbits[cur], (const int32_t *)nullptr);
      GG_HIP(hipGetLastError());
      bottom_up = true;
    } else if (bottom_up && nf < n / beta && nf < prev_nf) {
      TimedRegion t(g->prof, "bfs.bits_to_queue");
      GG_HIP(slot_resets_async(ctr, {{8 + cur, 0ull}}, s));
      hipLaunchKernelGGL(k_bits_to_queue, dim3(grid_for(words, 256, 4096)), dim3(256), 0, s, fbits[cur], words,
                         queue[cur], qsz[cur], (const int32_t *)nullptr);
      GG_HIP(hipGetLastError());
      bottom_up = false;
    }
    const int nxt = cur ^ 1;
    GG_HIP(slot_resets_async(ctr, {{0, 0ull}, {1, 0ull}, {2, 0ull}, {3, 0ull}, {8 + nxt, 0ull}}, s));
    double lvl_bytes;
    if (bottom_up) {
      {
        TimedRegion t(g->prof, "bfs.bottom_up");
        hipLaunchKernelGGL(k_bfs_bottom_up, dim3(bottom_up_grid(n)), dim3(kBuBlock), 0, s, g->dg.column_offsets,
                           g->dg.row_indices, first4, g->directed ? g->dg.row_offsets : nullptr, depth, visited,
                           fbits[cur], fbits[nxt], n, levels + 1, ctr, (const int32_t *)nullptr,
                           (const int32_t *)nullptr);
        GG_HIP(hipGetLastError());
      }
      GG_TRY(read_counters(g, 4));
      const int64_t N = g->h64()[0], E = g->h64()[2], U = g->h64()[3];
      lvl_bytes = 8.0 * U + 4.0 * E + n / 8.0 + 8.0 * N;
      g->prof.add_launch_bytes("bfs.bottom_up", lvl_bytes);
      examined_total += E;
      prev_nf = nf;
      nf = N;
      pull_levels++;
    } else {
      BfsPushOp op{g->dg.row_offsets, g->dg.column_indices, depth, visited, ctr, levels + 1, nullptr, 0ull};
      const bool small = nf <= kTierItems && mf <= kTierEdges;
      if (small) {  // as k_bfs_ctrl: a small level is one tiered launch
        TimedRegion t(g->prof, "bfs.top_down_small");
        hipLaunchKernelGGL((k_tiered_expand<BfsPushOp, false>), dim3(grid_for(kTierItems, kExpandBlock)),
                           dim3(kExpandBlock), 0, s, op, queue[cur], (const int32_t *)qsz[cur], queue[nxt], qsz[nxt],
                           g->opt.wave_degree, g->opt.block_degree, lb.edge_counter);
      } else {
        lbs_launch(op, false, queue[cur], qsz[cur], nf, (int64_t)n + g->m, lb, queue[nxt], qsz[nxt], s,
                   lbs_timer);
      }
      GG_HIP(hipGetLastError());
      GG_TRY(read_counters(g, 10));
      const int64_t E = g->h64()[2];
      const int64_t N = g->h_pinned[(8 + nxt) * 2];
      lvl_bytes = 12.0 * nf + 8.0 * E + 8.0 * N;
      if (small) {
        g->prof.add_launch_bytes("bfs.top_down_small", lvl_bytes);
      } else {
        g->prof.add_launch_bytes("bfs.lbs_count", 12.0 * nf);
        g->prof.add_launch_bytes("bfs.top_down", 8.0 * E + 8.0 * N);
      }
      examined_total += E;
      prev_nf = nf;
      nf = N;
    }
    bytes += lvl_bytes;
    mf = g->h64()[1];
    mu -= mf;
    reached += nf;
    reached_edges += mf;
    levels++;
    cur = nxt;
  }
  if (!timed) {
    GG_HIP(hipEventRecord(e1, s));
    GG_HIP(hipEventSynchronize(e1));
  }
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  finish_profile(g);
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->iterations = levels;
    stats->pull_iterations = pull_levels;
    stats->reached_vertices = reached;
    stats->reached_edges = reached_edges;
    stats->edges_inspected = examined_total;
    stats->alg_bytes = bytes;
    stats->gpu_ms = ms;
  }
  if (flags & GG_OUT_DEVICE) return GG_OK;
  return copy_out(g, depth_out, depth, 4 * (size_t)n, flags);
}

}  // namespace gg
