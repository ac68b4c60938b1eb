// Connected components by label propagation, device-driven (gg_cc default).
//
// Same semantics as the reference's CC program (Algorithms/CC/cc.h): every
// vertex starts with its own id as label and active (cc.h:68-71,118-140); an
// active vertex u offers label[u] to each neighbour v with label[v] > label[u]
// (cc.h:293-299), taken with atomicMin (cc.h:339); the vertices whose label
// dropped form the next frontier.  The fixed point is the smallest vertex id
// of each component, checked against CPUCC (cc.cu:127-169).
//
// What differs is the schedule.
//  * Direction: while the frontier holds more than n / kPullDiv vertices
//    (label propagation keeps most of a lattice active for a hundred
//    iterations) an iteration is a pull: every vertex takes the minimum label
//    over its neighbours (merge-path tiles of vertices + edges, so hub rows
//    spread over the whole grid; LDS minimum per row; a row inside one tile
//    is stored plainly, one spanning tiles with atomicMin) and the vertices
//    that changed are marked in the next frontier bitmap, counted once per
//    workgroup.  Smaller frontiers push with atomicMin from a queue; a pull
//    followed by a push first compacts the bitmap into that queue.  After a pull every edge (u, v) has label[v] <= the
//    label[u] it read, and a u that changed during the pull is in the next
//    frontier, so "every vertex whose label has not been offered to its
//    neighbours is in the frontier" holds across both kinds of step and the
//    fixed point is unchanged.
//  * Device-driven loop: an iteration is a fixed launch sequence (push, pull,
//    swap) whose kernels read queue pointers, sizes and the direction from a
//    device-side state block, so kIters iterations are captured once into a
//    hipGraph and replayed until the frontier is empty, one host check per
//    replay.  Bitmaps: a push clears, per frontier vertex, its word of the
//    bitmap its frontier was deduplicated with (so that bitmap is clean when
//    it becomes the next output bitmap); before a pull, and before the first
//    push after a pull, the output bitmap is cleared whole.
//
// Byte model: push iteration 16 per frontier vertex (id, offsets, label,
// bitmap word) + 8 per edge (column id, label probe); pull iteration
// 12 n + n / 8 (offsets, labels, bitmap) + 8 m (column ids, label gathers);
// both 8 per new frontier entry (label write, queue append).
#include <hip/hip_runtime.h>

#include <cstring>

#include "internal.h"

namespace gg {
namespace {

using namespace GASengine;
constexpr int kIters = 16;  // iterations per graph replay
constexpr int kBlock = 256;
constexpr int kPullDiv = 8;  // pull while the frontier exceeds n / kPullDiv
constexpr int kCcStride = 16;  // 128 bytes between the stripes of a counter

struct CcState {
  int32_t *q_cur, *q_nxt;
  uint32_t *b_cur, *b_nxt;  // dedup bitmaps of q_cur / q_nxt
  int32_t cur_size, nxt_size;
  int32_t iters, done;
  unsigned long long edges;
  unsigned long long frontier_total;  // sum of the frontier sizes consumed
  unsigned long long push_frontier;   // the same over push iterations
  int32_t pull, pulls;                // this iteration pulls; pull iterations so far
  int32_t prev_pull, pull_skip;       // last iteration pulled; pull kernel off (done || !pull)
  int32_t fill;                       // queue fill position of the bitmap compaction
  unsigned long long pull_changed;    // vertices changed by this iteration's pull
  // Striped per-workgroup sums (GASengine::block_counter_add, stride
  // kCcStride), folded into edges / pull_changed by k_lp_swap.
  unsigned long long edge_stripes[kStripes * kCcStride];
  unsigned long long changed_stripes[kStripes * kCcStride];
};

__global__ void k_lp_state_init(CcState *st, int32_t *q0, int32_t *q1, uint32_t *b0, uint32_t *b1, int32_t n) {
  st->q_cur = q0;
  st->q_nxt = q1;
  st->b_cur = b0;
  st->b_nxt = b1;
  st->cur_size = n;
  st->nxt_size = 0;
  st->iters = 0;
  st->done = n == 0;
  st->edges = 0;
  st->frontier_total = 0;
  st->push_frontier = 0;
  st->pull = (int64_t)n > (int64_t)n / kPullDiv;  // every vertex starts active
  st->pulls = 0;
  st->prev_pull = 0;
  st->pull_skip = !st->pull || st->done;
  st->fill = 0;
  st->pull_changed = 0;
  for (int q = 0; q < kStripes * kCcStride; q++) st->edge_stripes[q] = st->changed_stripes[q] = 0;
}

__device__ __forceinline__ bool offer(int32_t *label, int32_t v, int32_t lu, uint32_t *bits) {
  if (lu >= label[v]) return false;  // plain pre-check: a stale value is only larger
  const int32_t old = atomicMin(&label[v], lu);
  return lu < old && bit_test_and_set(bits, v);
}

// Frontier vertices with degree < wave_degree: one lane walks its row; longer
// rows: the whole wave.  New frontier vertices are staged in LDS and appended
// with one atomic per workgroup step.
__global__ void __launch_bounds__(kBlock)
k_lp_expand(CcState *st, const int32_t *__restrict__ ro, const int32_t *__restrict__ col, int32_t *label,
            int32_t wave_degree) {
  __shared__ int32_t s_q[kBlock * 4];
  __shared__ int32_t s_n, s_base;
  if (st->done || st->pull) return;
  const int32_t nf = st->cur_size;
  const int32_t *q = st->q_cur;
  uint32_t *bc = st->b_cur, *bn = st->b_nxt;
  const int lane = lane_id();
  unsigned long long edges = 0;
  for (int64_t base = (int64_t)blockIdx.x * kBlock; base < nf; base += (int64_t)gridDim.x * kBlock) {
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const int64_t i = base + threadIdx.x;
    int32_t u = -1, b = 0, e = 0, lu = 0;
    if (i < nf) {
      u = q[i];
      bc[u >> 5] = 0u;  // every set bit of that word belongs to a vertex of this frontier
      b = ro[u];
      e = ro[u + 1];
      lu = label[u];
      edges += (unsigned long long)(e - b);
    }
    const bool big = e - b >= wave_degree;
    // Short rows: lane-serial.  At most 4 appends per lane per step fit s_q;
    // beyond that, append straight to the global queue.
    if (!big)
      for (int32_t k = b; k < e; k++) {
        const int32_t v = col[k];
        if (offer(label, v, lu, bn)) {
          const int32_t at = atomicAdd(&s_n, 1);
          if (at < kBlock * 4)
            s_q[at] = v;
          else
            st->q_nxt[atomicAdd(&st->nxt_size, 1)] = v;
        }
      }
    // Long rows: the whole wave, one row at a time.
    uint64_t mask = ballot(big);
    while (mask) {
      const int l = __builtin_ctzll(mask);
      mask &= mask - 1;
      const int32_t wb = __shfl(b, l), we = __shfl(e, l), wl = __shfl(lu, l);
      for (int32_t k = wb + lane; k < we; k += kWave) {
        const int32_t v = col[k];
        if (offer(label, v, wl, bn)) {
          const int32_t at = atomicAdd(&s_n, 1);
          if (at < kBlock * 4)
            s_q[at] = v;
          else
            st->q_nxt[atomicAdd(&st->nxt_size, 1)] = v;
        }
      }
    }
    __syncthreads();
    const int32_t cnt = min(s_n, kBlock * 4);
    if (cnt) {
      if (threadIdx.x == 0) s_base = atomicAdd(&st->nxt_size, cnt);
      __syncthreads();
      for (int32_t k = threadIdx.x; k < cnt; k += kBlock) st->q_nxt[s_base + k] = s_q[k];
    }
    __syncthreads();
  }
  block_counter_add<kBlock, kCcStride>(st->edge_stripes, edges);
}

// Pull operator for GASengine::k_pull_staged over the (symmetric) adjacency.
// Runs on GASengine::k_pull_staged (neighbour ids load with the tile, their
// labels after; label[v] is prefetched).
struct CcPullOp {
  const int32_t *nbr;
  int32_t *label;
  CcState *st;                    // output bitmap st->b_nxt (cleared before the pull), count st->pull_changed
  unsigned long long count;

  struct Edge {
    int32_t u;
  };
  __device__ __forceinline__ Edge load_edge(int32_t slot) { return {nbr[slot]}; }
  __device__ __forceinline__ int32_t gather(const Edge &e) { return label[e.u]; }
  __device__ __forceinline__ void combine(int32_t *acc, int32_t val) {
    if (val < *acc) atomicMin(acc, val);
  }
  __device__ __forceinline__ int32_t vertex_state(int32_t v) { return label[v]; }
  // A row starts from its own current value: only smaller gathers touch LDS.
  __device__ __forceinline__ int32_t acc_start(int32_t state) { return state; }
  // k_pull_staged calls this with consecutive vertices on consecutive lanes
  // (v = first + lane for the active lanes), so the changed marks of a wave
  // are OR-ed into each bitmap word with one atomic; only bits the atomic
  // newly set are counted (a row spanning two tiles can be marked twice).
  // cur = label[v] at the start of the tile: exact when `whole`, otherwise
  // possibly stale (only ever larger), and the atomicMin decides.
  __device__ __forceinline__ void item_done(int32_t v, int32_t acc, bool whole, int32_t cur) {
    bool changed = false;
    if (acc < cur) {
      if (whole)
        label[v] = acc, changed = true;  // all of v's edges are in this tile: the only writer
      else
        changed = acc < atomicMin(&label[v], acc);
    }
    const uint64_t m = ballot(changed);
    if (!changed) return;
    const int lane = lane_id();
    const int64_t v0 = (int64_t)v - lane;  // vertex of lane 0
    const int64_t w = v >> 5;
    const int lo = (int)max<int64_t>(0, w * 32 - v0), hi = (int)min<int64_t>(64, w * 32 + 32 - v0);
    const uint64_t group = (m >> lo) & (hi - lo >= 64 ? ~0ull : ((1ull << (hi - lo)) - 1));
    if (lane != lo + __builtin_ctzll(group)) return;  // lowest changed lane of this word writes
    const uint32_t word = (uint32_t)(group << ((v0 + lo) & 31));
    const uint32_t old = atomicOr(&st->b_nxt[w], word);
    count += (unsigned long long)__popc(word & ~old);
  }
  __device__ __forceinline__ void finish() { block_counter_add<kExpandBlock, kCcStride>(st->changed_stripes, count); }
};

// Start of an iteration: clear the output bitmap before a pull, and before
// the first push after a pull, when also the pull's bitmap becomes the queue.
__global__ void __launch_bounds__(kBlock)
k_lp_prep(CcState *st, int32_t words) {
  __shared__ int32_t s_scan[kBlock / kWave + 1];
  __shared__ int32_t s_base;
  if (st->done || !(st->pull || st->prev_pull)) return;
  uint32_t *bn = st->b_nxt;
  const uint32_t *bc = st->b_cur;
  const bool compact = !st->pull;
  for (int64_t base = (int64_t)blockIdx.x * kBlock; base < words; base += (int64_t)gridDim.x * kBlock) {
    const int64_t w = base + threadIdx.x;
    uint32_t bits = 0u;
    if (w < words) {
      bn[w] = 0u;
      if (compact) bits = bc[w];
    }
    if (!compact) continue;
    int32_t total;
    const int32_t off = block_exclusive_scan<kBlock>(__popc(bits), s_scan, &total);
    if (total) {
      if (threadIdx.x == 0) s_base = atomicAdd(&st->fill, total);
      __syncthreads();
      int32_t at = s_base + off;
      while (bits) {
        const int b = __builtin_ctz(bits);
        bits &= bits - 1;
        st->q_cur[at++] = (int32_t)(w * 32 + b);
      }
      __syncthreads();
    }
  }
}

__global__ void k_lp_swap(CcState *st, int32_t n, int64_t m, int32_t max_iter) {
  if (st->done) return;
  unsigned long long pushed = 0, changed = 0;
  for (int q = 0; q < kStripes; q++) {
    pushed += st->edge_stripes[q * kCcStride];
    changed += st->changed_stripes[q * kCcStride];
    st->edge_stripes[q * kCcStride] = 0;
    st->changed_stripes[q * kCcStride] = 0;
  }
  st->edges += pushed;
  st->pull_changed += changed;
  if (st->pull) {
    st->edges += (unsigned long long)m;
    st->pulls++;
    st->nxt_size = (int32_t)st->pull_changed;
    st->pull_changed = 0;
  } else {
    st->push_frontier += (unsigned long long)st->cur_size;
  }
  st->prev_pull = st->pull;
  st->fill = 0;
  int32_t *q = st->q_cur;
  st->q_cur = st->q_nxt;
  st->q_nxt = q;
  uint32_t *b = st->b_cur;
  st->b_cur = st->b_nxt;
  st->b_nxt = b;
  st->frontier_total += (unsigned long long)st->cur_size;
  st->cur_size = st->nxt_size;
  st->nxt_size = 0;
  st->iters++;
  st->pull = st->cur_size > n / kPullDiv;
  if (st->cur_size == 0 || st->iters >= max_iter) st->done = 1;
  st->pull_skip = !st->pull || st->done;
}

}  // namespace

int cc_label_propagation(gg_graph *g, int32_t *label_out, uint32_t flags, gg_stats *stats) {
  GG_TRY(activate(g));
  if (g->directed) return set_error(GG_ERR_UNSUPPORTED, "CC needs a symmetric (undirected) graph");
  const int32_t n = g->n;
  const size_t words = ((size_t)n + 31) / 32;
  hipStream_t s = g->stream;
  int32_t *label;
  if (flags & GG_OUT_DEVICE) {
    GG_CHECK_ARG(label_out, "label is NULL");
    label = label_out;
  } else {
    GG_TRY(g->vtx[1].ensure(4 * (size_t)n));
    label = g->vtx[1].as<int32_t>();
  }
  GG_TRY(g->frontier[0].ensure(4 * (size_t)n));
  GG_TRY(g->frontier[1].ensure(4 * (size_t)n));
  GG_TRY(g->bits[0].ensure(4 * words));
  GG_TRY(g->bits[1].ensure(4 * words));
  GG_TRY(g->batch[7].ensure(sizeof(CcState)));
  CcState *st = g->batch[7].as<CcState>();
  int32_t *q0 = g->frontier[0].as<int32_t>(), *q1 = g->frontier[1].as<int32_t>();
  uint32_t *b0 = g->bits[0].as<uint32_t>(), *b1 = g->bits[1].as<uint32_t>();
  const int32_t max_iter = g->opt.iter_num;
  const int grid = g->opt.expand_grid;
  g->prof.enabled = (flags & GG_INSTRUMENT) != 0;

  ScopedEvents ev;
  GG_HIP(ev.create());
  hipEvent_t e0 = ev.e0, e1 = ev.e1;
  GG_HIP(hipEventRecord(e0, s));
  {
    TimedRegion t(g->prof, "cc.init");
    GG_HIP(sequence_async<int32_t>(label, n, s));
    GG_HIP(sequence_async<int32_t>(q0, n, s));  // every vertex starts active
    GG_HIP(fill_async<uint32_t>(b0, 0u, (int64_t)words, s));
    GG_HIP(fill_async<uint32_t>(b1, 0u, (int64_t)words, s));
    hipLaunchKernelGGL(k_lp_state_init, dim3(1), dim3(1), 0, s, st, q0, q1, b0, b1, n);
    GG_HIP(hipGetLastError());
  }
  g->prof.add_bytes("cc.init", 12.0 * n + words * 8.0);
  const int32_t *parts = nullptr;
  GG_TRY(pull_partition(g, &parts));  // merge-path tiles over (vertices, adjacency), cached per graph
  CcPullOp pull_op{g->dg.column_indices, label, st, 0ull};
  ScopedGraph cg;
  GG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  for (int it = 0; it < kIters; it++) {
    hipLaunchKernelGGL(k_lp_prep, dim3(grid_for((int64_t)words, kBlock, 1024)), dim3(kBlock), 0, s, st,
                       (int32_t)words);
    hipLaunchKernelGGL(k_lp_expand, dim3(grid), dim3(kBlock), 0, s, st, g->dg.row_offsets, g->dg.column_indices,
                       label, g->opt.wave_degree);
    hipLaunchKernelGGL((k_pull_staged<CcPullOp, int32_t>), dim3(grid), dim3(kExpandBlock), 0, s, pull_op,
                       g->dg.row_offsets, n, g->m, parts, &st->pull_skip);
    hipLaunchKernelGGL(k_lp_swap, dim3(1), dim3(1), 0, s, st, n, g->m, max_iter);
  }
  hipError_t cap = hipStreamEndCapture(s, &cg.graph);
  if (cap != hipSuccess) return set_error(GG_ERR_HIP, "graph capture failed: %s", hipGetErrorString(cap));
  GG_HIP(hipGraphInstantiate(&cg.exec, cg.graph, nullptr, nullptr, 0));
  struct H {
    int32_t iters, done;
    unsigned long long edges, frontier_total, push_frontier;
    int32_t pull, pulls;
  };
  H h{};
  int rc = GG_OK;
  for (;;) {
    {
      TimedRegion t(g->prof, "cc.labelprop");
      if (hipGraphLaunch(cg.exec, s) != hipSuccess) {
        rc = set_error(GG_ERR_HIP, "graph launch failed");
        break;
      }
    }
    if (hipMemcpyAsync(g->h_pinned, &st->iters, sizeof(H), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      rc = set_error(GG_ERR_HIP, "state readback failed");
      break;
    }
    std::memcpy(&h, g->h_pinned, sizeof(H));
    if (h.done) break;
  }
  GG_TRY(rc);
  GG_HIP(hipEventRecord(e1, s));
  GG_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  const double F = (double)h.frontier_total;
  const double bytes = 12.0 * n + words * 8.0 + 16.0 * (double)h.push_frontier + 8.0 * (double)h.edges +
                       (double)h.pulls * (12.0 * n + n / 8.0) + 8.0 * (F - n);
  g->prof.add_bytes("cc.labelprop", bytes - 12.0 * n - words * 8.0);
  finish_profile(g);
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->iterations = h.iters;
    stats->pull_iterations = h.pulls;
    stats->edges_inspected = (int64_t)h.edges;
    stats->alg_bytes = bytes;
    stats->gpu_ms = ms;
    GG_TRY(reach_stats(g, label, INT32_MIN, stats));
  }
  if (flags & GG_OUT_DEVICE) return GG_OK;
  return copy_out(g, label_out, label, 4 * (size_t)n, flags);
}

}  // namespace gg
