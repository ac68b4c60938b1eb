// SSSP by bucketed, direction-optimising relaxation (default path of gg_sssp).
//
// Output contract of the reference's SSSP program (Algorithms/SSSP/sssp.h;
// CPUSSSP sssp.cu:126-183): dist[src] = 0, unreached = 100000000, int32
// weights, a path length >= 100000000 never relaxes (CPUSSSP compares a
// 64-bit sum against the sentinel).  Shortest-path distances are the unique
// fixed point, so any schedule of monotone relaxations that runs to
// quiescence reproduces them bit for bit; the generic GAS path
// (GG_SSSP_GAS) keeps the reference's own schedule.
//
// State: dist[n], dirty[n] (byte: distance dropped since the vertex was last
// expanded), frontier queue.  Per iteration:
//   collect  every dirty vertex with dist < bound joins the frontier (queue,
//            dirty cleared), in vertex order, with its out-degree offset:
//            one launch that is also the LBS count and scan (a single-pass
//            scan across workgroups) and the controller (its
//            last workgroup decides push, tiered push, pull or end).  Empty
//            frontier with dirty vertices left: the next collect moves the
//            bound to the bucket of the smallest distance still dirty; when
//            the last push set no distance below the bound (its smallest is
//            kept in a counter) the collect moves the bound first and never
//            finds that empty frontier.
//   push     frontier edges <= m / pull_alpha: LBS partition + expand over
//            the queue
//            (one tiered launch instead when the frontier is small: at most
//            kTierItems vertices and kTierEdges edges),
//            nd = dist[v] + w, plain pre-check then atomicMin(dist[nb]); an
//            improvement stores dirty[nb] = 1 (no atomic, no enqueue).  On an
//            undirected graph whose in-edges are packed (below) the push
//            reads the same packed words: they are the CSR's own slots.
//   pull     otherwise, unless the previous iteration pulled: merge-path
//            pass over every vertex's in-edges, min
//            of dist[u] + w over all in-neighbours u reduced in LDS (each
//            thread's edges loaded back to back), written with a plain store
//            when the tile holds all of the vertex's in-edges (atomicMin at
//            tile seams).  A pull ends the buckets: the bound becomes
//            unlimited, so the next collects take every dirty vertex (the
//            pull relaxed every in-edge, so bucket order saves no more work
//            and the tail needs fewer iterations).
//
// Byte model: collect n + 16F (dirty bytes; dist and offsets read, queue
// slot and offset written per frontier vertex); push 16F + 12E + 8N;
// pull 4n + 12m (offsets; source, weight and distance per in-edge: the same
// 12 bytes per inspected edge as the push).  Both count 12 bytes per edge
// whichever layout is read; the packed words move 8 of them.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstring>
#include <vector>

#include "internal.h"

namespace gg {
namespace {

using namespace GASengine;
constexpr int32_t kInf = 100000000;
constexpr int kLoopIters = 12;  // iterations per replay after the first
static_assert(kLoopIters <= kSsspReplayMax, "replay lengths are cached up to kSsspReplayMax");

// counters (striped, GASengine::block_counter_add): [2] LBS edge counter,
// [4] improvements, [16..19] the collect's dirty-left counts and minimum
// distances (two pairs, by collect parity); slot 8 holds the int32 queue
// sizes the expand kernels append to (stripe 0 only; the push never
// appends).  Reached vertices
// and their degree sum (the run's statistics) come from one pass over the
// final distances (reach_stats), not from the relaxations.
// Small pushes run as one tiered launch instead of the LBS partition and expand.
constexpr int32_t kTierItems = 2048;
constexpr int64_t kTierEdges = 65536;
// Counter slots of the pull: distance gathers, then the in-edges and rows of
// the tiles it visited (slot 4: improvements).
constexpr int kGatheredSlot = 5;
constexpr int kStreamedEdgesSlot = 6;
constexpr int kStreamedRowsSlot = 7;

// Counter slot of the smallest distance an iteration's relaxations set
// (striped minimum; the collect that follows reads it, its last workgroup
// resets it).
constexpr int kRelaxMinSlot = 20;

struct Tally {
  unsigned long long improved = 0;
  int32_t set_min = INT_MAX;  // smallest distance this thread's improvements set
  __device__ __forceinline__ void add(int32_t d) {
    improved++;
    set_min = min(set_min, d);
  }
  // All threads of the workgroup must call it.
  __device__ __forceinline__ void flush(unsigned long long *counters) {
    block_counter_add<kExpandBlock>(&counters[4], improved);
    __shared__ int32_t s_min[kExpandBlock / kWave];
    int32_t mn = set_min;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mn = min(mn, __shfl_xor(mn, off));
    if (lane_id() == 0) s_min[threadIdx.x / kWave] = mn;
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w = 1; w < kExpandBlock / kWave; w++) mn = min(mn, s_min[w]);
      if (mn != INT_MAX) atomicMin(stripe_of(&counters[kRelaxMinSlot]), (unsigned long long)mn);
    }
  }
};

// PACKED: the CSR slot's target and weight come from one word of the packed
// in-edge array, which is the CSR's own slot order on an undirected graph
// whose in-edge weights are its out-edge weights: one coalesced 4-byte load
// per edge instead of two.
template <bool PACKED>
struct PushOpT {
  static constexpr bool kItemValues = true;
  const int32_t *ro;
  const int32_t *col;
  const int32_t *w;
  const uint32_t *pk;  // PACKED: (target << wbits) | weight per CSR slot
  int32_t wbits;
  uint32_t wmask;
  int32_t *dist;
  uint8_t *dirty;
  unsigned long long *counters;
  Tally tally;

  __device__ __forceinline__ void prologue(int32_t) {}
  __device__ __forceinline__ int32_t degree(int32_t v) { return ro[v + 1] - ro[v]; }
  __device__ __forceinline__ int32_t row_begin(int32_t v) { return ro[v]; }
  __device__ __forceinline__ int32_t item_value(int32_t v) { return dist[v]; }
  // Staged form: column and weight of all of a thread's edges are loaded
  // before the first distance probe.  The probe is a plain (cacheable) load
  // made just before the edge's atomic, so it sees the thread's own earlier
  // updates; a stale, larger value only lets an unneeded atomicMin through,
  // whose returned old value filters it.
  struct Loads {
    int32_t nb;
    int64_t nd;
  };
  __device__ __forceinline__ Loads load(int32_t, int32_t dv, int32_t slot) {
    if constexpr (PACKED) {
      const uint32_t e = pk[slot];
      return {(int32_t)(e >> wbits), (int64_t)dv + (int32_t)(e & wmask)};
    } else {
      return {col[slot], (int64_t)dv + w[slot]};
    }
  }
  __device__ __forceinline__ bool apply(int32_t, const Loads &l, int32_t &) {
    if (l.nd >= dist[l.nb]) return false;
    const int32_t nd = (int32_t)l.nd;
    const int32_t old = atomicMin(&dist[l.nb], nd);
    if (nd >= old) return false;
    tally.add(nd);
    dirty[l.nb] = 1;
    return false;
  }
  __device__ __forceinline__ bool edge(int32_t v, int32_t dv, int32_t slot, int32_t &out) {
    return apply(v, load(v, dv, slot), out);
  }
  __device__ __forceinline__ void finish() { tally.flush(counters); }
};
using PushOp = PushOpT<false>;
using PackedPushOp = PushOpT<true>;

// Relaxes every in-edge, not only those of frontier vertices: the pass reads
// all in-edges anyway, dist[u] + w is a path length whatever u's state, and
// distances lowered earlier in the same pass propagate at once.  dist[u] is a
// plain load (it may be stale, i.e. larger, and then only relaxes less).
// Runs on GASengine::k_pull_staged: an edge's source and weight load with
// the tile, the gather of dist[u] after, dist[v] is prefetched.
struct PullOp {
  const int32_t *ri;
  const int32_t *wc;
  int32_t *dist;
  uint8_t *dirty;
  unsigned long long *counters;
  Tally tally;
  const int32_t *settled;  // SsspCtl::settled: rows of vertices with dist <= *settled are skipped
  const int32_t *rows;     // k_pull_rows' prefix count of the rows that can still change
  unsigned long long gathered = 0, streamed_edges = 0, streamed_rows = 0;

  struct Edge {
    int32_t u, w;
  };
  __device__ __forceinline__ Edge load_edge(int32_t slot) {
    return {__builtin_nontemporal_load(&ri[slot]), __builtin_nontemporal_load(&wc[slot])};
  }
  __device__ __forceinline__ const int32_t *row_prefix() const { return rows; }
  __device__ __forceinline__ void tile_visited(int32_t edges, int32_t items) {
    streamed_edges += (unsigned long long)edges;
    streamed_rows += (unsigned long long)items;
  }
  __device__ __forceinline__ int32_t row_limit() const { return *settled; }
  __device__ __forceinline__ int32_t skipped() const { return INT_MAX; }
  __device__ __forceinline__ int32_t gather(const Edge &e) {
    gathered++;
    const int64_t c = (int64_t)dist[e.u] + e.w;
    return c < kInf ? (int32_t)c : INT_MAX;
  }
  __device__ __forceinline__ void combine(int32_t *acc, int32_t val) {
    if (val < *acc) atomicMin(acc, val);
  }
  __device__ __forceinline__ int32_t vertex_state(int32_t v) { return dist[v]; }
  // A row starts from its own current value: only smaller gathers touch LDS.
  __device__ __forceinline__ int32_t acc_start(int32_t state) { return state; }
  // cur: dist[v] at the start of the tile; exact when `whole`, otherwise
  // possibly stale (larger), and the atomicMin decides.
  __device__ __forceinline__ void item_done(int32_t v, int32_t acc, bool whole, int32_t cur) {
    if (acc == INT_MAX || acc >= cur) return;
    int32_t old = cur;
    if (whole) {
      dist[v] = acc;
    } else {
      old = atomicMin(&dist[v], acc);
      if (acc >= old) return;
    }
    tally.add(acc);
    dirty[v] = 1;
  }
  __device__ __forceinline__ void finish() {
    tally.flush(counters);
    block_counter_add<kExpandBlock>(&counters[kGatheredSlot], gathered);
    block_counter_add<kExpandBlock>(&counters[kStreamedEdgesSlot], streamed_edges);
    block_counter_add<kExpandBlock>(&counters[kStreamedRowsSlot], streamed_rows);
  }
};

// The pull over packed in-edges: one 32-bit word per edge holds the source
// (high bits) and the weight (low wbits bits), so an edge is one load of the
// stream instead of two.
struct PackedPullOp {
  const uint32_t *pk;
  int32_t wbits;
  uint32_t wmask;
  int32_t *dist;
  uint8_t *dirty;
  unsigned long long *counters;
  Tally tally;
  const int32_t *settled;  // SsspCtl::settled: rows of vertices with dist <= *settled are skipped
  const int32_t *rows;     // k_pull_rows' prefix count of the rows that can still change
  unsigned long long gathered = 0, streamed_edges = 0, streamed_rows = 0;

  struct Edge {
    int32_t u, w;
  };
  __device__ __forceinline__ const int32_t *row_prefix() const { return rows; }
  __device__ __forceinline__ void tile_visited(int32_t edges, int32_t items) {
    streamed_edges += (unsigned long long)edges;
    streamed_rows += (unsigned long long)items;
  }
  // The edge stream is read once per pull: a non-temporal load keeps it from
  // displacing the gathered distances in L2.
  __device__ __forceinline__ Edge load_edge(int32_t slot) {
    const uint32_t e = __builtin_nontemporal_load(&pk[slot]);
    return {(int32_t)(e >> wbits), (int32_t)(e & wmask)};
  }
  __device__ __forceinline__ int32_t row_limit() const { return *settled; }
  __device__ __forceinline__ int32_t skipped() const { return INT_MAX; }
  __device__ __forceinline__ int32_t gather(const Edge &e) {
    gathered++;
    const int64_t c = (int64_t)dist[e.u] + e.w;
    return c < kInf ? (int32_t)c : INT_MAX;
  }
  __device__ __forceinline__ void combine(int32_t *acc, int32_t val) {
    if (val < *acc) atomicMin(acc, val);
  }
  __device__ __forceinline__ int32_t vertex_state(int32_t v) { return dist[v]; }
  // A row starts from its own current value: only smaller gathers touch LDS.
  __device__ __forceinline__ int32_t acc_start(int32_t state) { return state; }
  __device__ __forceinline__ void item_done(int32_t v, int32_t acc, bool whole, int32_t cur) {
    if (acc == INT_MAX || acc >= cur) return;
    int32_t old = cur;
    if (whole) {
      dist[v] = acc;
    } else {
      old = atomicMin(&dist[v], acc);
      if (acc >= old) return;
    }
    tally.add(acc);
    dirty[v] = 1;
  }
  __device__ __forceinline__ void finish() {
    tally.flush(counters);
    block_counter_add<kExpandBlock>(&counters[kGatheredSlot], gathered);
    block_counter_add<kExpandBlock>(&counters[kStreamedEdgesSlot], streamed_edges);
    block_counter_add<kExpandBlock>(&counters[kStreamedRowsSlot], streamed_rows);
  }
};

__global__ void k_pack_in_edges(const int32_t *__restrict__ src, const int32_t *__restrict__ w, int64_t m,
                                int32_t wbits, uint32_t *__restrict__ pk) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += (int64_t)gridDim.x * blockDim.x)
    pk[k] = ((uint32_t)src[k] << wbits) | (uint32_t)w[k];
}

int bits_for(int64_t x) {  // bits to hold 0..x
  int b = 1;
  while (b < 40 && (((int64_t)1 << b) <= x)) b++;
  return b;
}

// Packed in-edges of the graph for weights wc (CSC order), rebuilt when the
// values change; *pk = nullptr when a weight is negative or source id and
// weight do not fit one 32-bit word together.
int packed_in_edges(gg_graph *g, const int32_t *wc, const uint32_t **pk, int32_t *wbits) {
  *pk = nullptr;
  *wbits = 0;
  if (g->m == 0) return GG_OK;
  if (g->pk_version != g->values_version) {
    g->pk_version = g->values_version;
    g->pk_wbits = 0;
    hipStream_t s = g->stream;
    // The CSC-ordered weights are a permutation of the CSR values: same range.
    int32_t h[2];
    GG_TRY(weight_range(g, &h[0], &h[1]));
    const int wb = bits_for(h[1]), ub = bits_for((int64_t)g->n - 1);
    if (h[0] >= 0 && wb + ub <= 32) {
      GG_TRY(g->pk_edges.ensure(4 * (size_t)g->m));
      hipLaunchKernelGGL(k_pack_in_edges, dim3(grid_for(g->m, 256, 16384)), dim3(256), 0, s, g->dg.row_indices, wc,
                         g->m, wb, g->pk_edges.as<uint32_t>());
      GG_HIP(hipGetLastError());
      g->pk_wbits = wb;
    } else {
      g->pk_edges.reset();
    }
  }
  if (g->pk_wbits > 0) {
    *pk = g->pk_edges.as<uint32_t>();
    *wbits = g->pk_wbits;
  }
  return GG_OK;
}


// Counter slots of the collect (striped): dirty vertices left above the
// bound and their smallest distance, one pair per collect parity (slots
// kLeftSlot + 2p and kLeftSlot + 2p + 1).  Collect e adds into pair e & 1 and
// reads pair (e - 1) & 1 when it has to move the bound; its last workgroup
// clears the pair it read.
constexpr int kLeftSlot = 16;

// State of a run, on the device.  Every collect reads it at the start of
// each workgroup; only the collect's last workgroup writes it, after its
// scan has read every other workgroup's published aggregate (so every
// workgroup has read the state by then), and the push / pull launches after
// the collect read its decisions.
struct SsspCtl {
  int32_t bound;      // frontier bound of the current bucket
  int32_t push_size;  // frontier items the LBS push sees: the queue size when pushing by LBS, else 0
  int32_t tier_size;  // frontier items the tiered push sees: the queue size when the push is small, else 0
  int32_t pull_skip;  // 0 when this iteration pulls
  int32_t done;
  int32_t advance;    // the last collect selected nothing: the next one moves the bound (or ends the run)
  int32_t epoch;      // collects run; its parity picks the scan words and counter slots
  int32_t iters, pulls;
  int32_t steps;      // collects that ran (the last one ends the run)
  int32_t last_nf;    // frontier size of the last collect
  int32_t prev_pull;  // the last iteration pulled
  int32_t settled;    // a pull skips the in-rows of vertices with dist <= settled (-1: none)
  int32_t stall;      // a scheduled replay's collect decided a kind its slot lacks (SsspKind, else 0)
  long long last_deg; // its out-degree sum
  long long edges;    // edges inspected
  double bytes;       // algorithmic bytes (model in the header comment)
  uint8_t kinds[kSpecMax];  // kind of each collect's iteration (SsspKind), the first kSpecMax collects
};

// Kinds of an iteration, as its collect decides (what its slot in a
// scheduled replay launches); bit k of a slot's mask allows kind k.  A
// collect that only moves the bound or ends the run is kind 0, which every
// slot allows.
enum SsspKind : int {
  kSsspCollectOnly = 0,
  kSsspLbs = 1,   // LBS partition + expand
  kSsspTier = 2,  // tiered push
  kSsspPull = 3,  // row prefix + pull
};
constexpr int32_t kSsspMaskAll = 0xf;
static_assert(sizeof(SsspCtl) % 8 == 0, "read back in 8-byte words");

// A workgroup takes kCollectStep consecutive vertices per tile, 16 per
// thread: the thread's 16 dirty bytes arrive in one 16-byte load (the dirty
// array is padded with zeros to a multiple of 16), and a tile where no byte is
// set ends there, after one barrier.  For the set bytes the thread reads the
// 16 distances and 17 row offsets as 16-byte vectors (VEC: dist and the
// offsets 16-byte aligned and the 16 vertices inside n; otherwise one load
// per set byte).
constexpr int kCollectBlock = 512;
constexpr int kCollectPer = 16;
constexpr int kCollectStep = kCollectPer * kCollectBlock;
// At most this many workgroups (each then takes a run of consecutive tiles),
// so that a workgroup's threads read all its predecessors' aggregates at once.
constexpr int kCollectGridMax = kCollectBlock;
// Frontier size and out-degree sum share one scan word: 31 bits each.
constexpr int kPackShift = 31;
constexpr uint64_t kPackMask = (1ull << kPackShift) - 1;

// Bits j of the result: byte j of the 16 bytes is nonzero.
__device__ __forceinline__ uint32_t nonzero_bytes(uint4 d) {
  const uint32_t w[4] = {d.x, d.y, d.z, d.w};
  uint32_t m = 0;
#pragma unroll
  for (int j = 0; j < 16; j++) m |= ((w[j >> 2] >> (8 * (j & 3))) & 0xffu) ? (1u << j) : 0u;
  return m;
}

// The 16 bytes with byte j cleared for every bit j of `clear`.
__device__ __forceinline__ uint4 clear_bytes(uint4 d, uint32_t clear) {
  uint32_t w[4] = {d.x, d.y, d.z, d.w};
#pragma unroll
  for (int j = 0; j < 16; j++)
    if (clear & (1u << j)) w[j >> 2] &= ~(0xffu << (8 * (j & 3)));
  return make_uint4(w[0], w[1], w[2], w[3]);
}

struct CollectArgs {
  uint8_t *dirty;
  const int32_t *dist;
  const int32_t *ro;
  int32_t n;
  int32_t tiles;            // kCollectStep-vertex tiles over n
  int32_t tiles_per_block;  // consecutive tiles per workgroup
  int32_t max_iter;
  int32_t *queue;
  int32_t *scan_local;      // LBS offsets: global exclusive out-degree prefix per queue slot
  int32_t *partials;        // LBS partials: 0 (scan_local holds the global prefix)
  LbsInfo *info;
  unsigned long long *ctr;
  unsigned long long *edge_counter;
  uint64_t *status;         // scan words, 2 x gridDim.x (one set per collect parity)
  SsspCtl *ctl;
  int64_t m, delta, pull_alpha;
  double pull_bytes;
  int32_t mask;             // kinds this collect's slot launches kernels for (kSsspMaskAll: all)
};

// The thread's 16 vertices from v0: bit j of the result selects v0 + j
// (dirty and dist < bound), deg[j] its out-degree.  ACC: count the dirty
// vertices left (dist >= bound) and their smallest distance.
template <bool VEC, bool ACC>
__device__ __forceinline__ uint32_t collect_select(const CollectArgs &a, int64_t v0, int32_t bound, uint4 &d4,
                                                   int32_t (&deg)[kCollectPer], unsigned long long &left,
                                                   int32_t &dmin) {
  d4 = make_uint4(0u, 0u, 0u, 0u);
  if (v0 < a.n) d4 = *reinterpret_cast<const uint4 *>(a.dirty + v0);
  const uint32_t nz = nonzero_bytes(d4);
  uint32_t sel = 0;
#pragma unroll
  for (int j = 0; j < kCollectPer; j++) deg[j] = 0;
  if (!nz) return 0;
  if (VEC && v0 + kCollectPer <= a.n) {
    int32_t dv[16], rv[17];
    const int4 *dp = reinterpret_cast<const int4 *>(a.dist + v0);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int4 x = dp[q];
      dv[4 * q] = x.x;
      dv[4 * q + 1] = x.y;
      dv[4 * q + 2] = x.z;
      dv[4 * q + 3] = x.w;
    }
    const int4 *rp = reinterpret_cast<const int4 *>(a.ro + v0);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int4 x = rp[q];
      rv[4 * q] = x.x;
      rv[4 * q + 1] = x.y;
      rv[4 * q + 2] = x.z;
      rv[4 * q + 3] = x.w;
    }
    rv[16] = a.ro[v0 + 16];  // v0 + 16 <= n: offsets has n + 1 entries
#pragma unroll
    for (int j = 0; j < 16; j++) {
      if (!(nz & (1u << j))) continue;
      if (dv[j] < bound) {
        sel |= 1u << j;
        deg[j] = rv[j + 1] - rv[j];
      } else if (ACC) {
        left++;
        dmin = min(dmin, dv[j]);
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < 16; j++) {
      if (!(nz & (1u << j))) continue;
      const int64_t v = v0 + j;  // a set byte lies below n
      const int32_t d = a.dist[v];
      if (d < bound) {
        sel |= 1u << j;
        deg[j] = a.ro[v + 1] - a.ro[v];
      } else if (ACC) {
        left++;
        dmin = min(dmin, d);
      }
    }
  }
  return sel;
}

// Writes the thread's selected vertices from queue slot *pos on (global
// out-degree offset *eoff on) and clears their dirty bytes.
__device__ __forceinline__ void collect_write(const CollectArgs &a, int64_t v0, uint4 d4, uint32_t sel,
                                              const int32_t (&deg)[kCollectPer], int32_t &pos, int32_t &eoff) {
  if (!sel) return;
#pragma unroll
  for (int j = 0; j < kCollectPer; j++) {
    if (!(sel & (1u << j))) continue;
    a.queue[pos] = (int32_t)(v0 + j);
    a.scan_local[pos] = eoff;
    if (pos % kCountTile == 0) a.partials[pos / kCountTile] = 0;
    pos++;
    eoff += deg[j];
  }
  *reinterpret_cast<uint4 *>(a.dirty + v0) = clear_bytes(d4, sel);
}

// The decisions of one iteration, taken by the collect's last workgroup
// once the frontier size nf and its out-degree sum are known: bucket
// advance, push (LBS or tiered) or pull, end.
__device__ void sssp_decide(const CollectArgs &a, int32_t bound, int32_t nf, int64_t degsum) {
  SsspCtl *c = a.ctl;
  if (c->steps < kSpecMax) c->kinds[c->steps] = kSsspCollectOnly;
  c->epoch++;
  c->steps++;
  c->bound = bound;
  c->advance = 0;
  c->push_size = 0;
  c->tier_size = 0;
  c->pull_skip = 1;
  c->last_nf = nf;
  c->last_deg = degsum;
  a.info->edges = 0;
  a.info->path = 0;
  c->bytes += a.n + 16.0 * nf;
  if (c->iters >= a.max_iter) {
    c->done = 1;
    return;
  }
  if (nf == 0) {
    // Nothing selected below the bound.  With the bound at kInf every dirty
    // vertex was selectable (a dirty vertex has dist < kInf), so the run is
    // over; otherwise the next collect moves the bound or finds it over.
    if (bound == kInf)
      c->done = 1;
    else
      c->advance = 1;
    return;
  }
  // Pull when the frontier's edges exceed m / pull_alpha, but never twice in
  // a row: after a pull, a push costs about as much per edge as a pull and
  // touches only the frontier's edges (degsum <= m).
  const bool pull = a.pull_alpha > 0 && !c->prev_pull && degsum > a.m / a.pull_alpha;
  c->prev_pull = pull;
  if (pull) {
    c->pull_skip = 0;
    // Every dirty vertex has dist >= bound - delta (the bucket's lower edge:
    // relaxations from the bucket's vertices cannot go below it, weights
    // being >= 0), so a vertex at or below that edge cannot improve: its
    // in-row needs no gathers in this pull.  Once the buckets have ended
    // (bound at kInf) the last such edge still holds: every later
    // relaxation starts from a dirty vertex at or above it.
    if (bound < kInf) c->settled = (int32_t)max<int64_t>((int64_t)bound - a.delta, -1);
    c->bound = kInf;  // after a pull, buckets end: every dirty vertex is collected
    c->bytes += a.pull_bytes;  // k_pull_rows; the pull's own bytes and edges are counted as it runs
    c->pulls++;
  } else {
    if (nf <= kTierItems && degsum <= kTierEdges) {
      c->tier_size = nf;
    } else {
      c->push_size = nf;
      a.info->edges = (int32_t)degsum;
      a.info->path = (int64_t)nf + degsum;
      atomicAdd(a.edge_counter, (unsigned long long)degsum);
    }
    c->bytes += 16.0 * nf + 12.0 * (double)degsum;
    c->edges += degsum;
  }
  c->iters++;
  const int kind = pull ? kSsspPull : c->tier_size > 0 ? kSsspTier : kSsspLbs;
  if (c->steps <= kSpecMax) c->kinds[c->steps - 1] = (uint8_t)kind;
  if (!((a.mask >> kind) & 1)) {
    // The slot lacks this kind's kernels: the iteration is decided and its
    // frontier collected, but its kernels are held back (every gate closed:
    // the LBS expand runs over info->path) until k_sssp_resume opens them
    // outside this replay.
    c->stall = kind;
    c->push_size = 0;
    c->tier_size = 0;
    c->pull_skip = 1;
    a.info->edges = 0;
    a.info->path = 0;
  }
}

// After a stalled replay: opens the held-back iteration's gates (the
// collect's decision, from the frontier size and degree sum it recorded).
__global__ void k_sssp_resume(SsspCtl *c, LbsInfo *info) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int k = c->stall;
  c->stall = 0;
  if (k == kSsspLbs) {
    c->push_size = c->last_nf;
    info->edges = (int32_t)c->last_deg;
    info->path = (int64_t)c->last_nf + c->last_deg;
  } else if (k == kSsspTier) {
    c->tier_size = c->last_nf;
  } else if (k == kSsspPull) {
    c->pull_skip = 0;
  }
}

// The collect, fused with the LBS count and scan and with the controller:
// every dirty vertex with dist < bound joins the frontier (dirty cleared),
// at a queue slot and out-degree offset from a single-pass scan across the
// workgroups, so the queue comes out in vertex order with the LBS
// offsets the partition and expand read (partials all 0).  The last
// workgroup then takes the iteration's decisions.  When the previous collect
// selected nothing, every workgroup first moves the bound to the bucket of
// the smallest distance still dirty (the same value everywhere), or finds the
// run finished.
template <bool VEC>
__global__ void __launch_bounds__(kCollectBlock) k_sssp_collect(CollectArgs a) {
  SsspCtl *c = a.ctl;
  if (c->done || c->stall) return;
  const int par = c->epoch & 1;
  int32_t bound = c->bound;
  const int64_t G = gridDim.x, b = blockIdx.x;
  const bool last = b == G - 1;
  // The smallest distance the last push set (INT_MAX: it improved nothing).
  // Every dirty vertex is either left from the previous collect (at or
  // above the bound then) or was improved since, to at least that distance;
  // so when it is at or above the bound, nothing below the bound is dirty,
  // and the bound moves now (the collect that would find nothing is
  // skipped), to the bucket of the smallest dirty distance.  Not after a
  // pull, which is not followed by buckets.
  unsigned long long rmin = a.ctr[kRelaxMinSlot];
  for (int q = 1; q < kStripes; q++) rmin = min(rmin, a.ctr[kRelaxMinSlot + q * kCounterSlots]);
  const bool skip_empty = !c->advance && c->iters > 0 && !c->prev_pull && rmin >= (unsigned long long)bound;
  if (c->advance || skip_empty) {
    const unsigned long long *prev = a.ctr + kLeftSlot + 2 * (par ^ 1);
    const unsigned long long left = stripe_sum(prev);
    if (left == 0 && (c->advance || rmin == (unsigned long long)INT_MAX)) {  // nothing dirty anywhere: the run is over
      if (last && threadIdx.x == 0) {
        if (c->steps < kSpecMax) c->kinds[c->steps] = kSsspCollectOnly;
        c->steps++;
        c->done = 1;
      }
      return;
    }
    unsigned long long dm = prev[1];
    for (int q = 1; q < kStripes; q++) dm = min(dm, prev[1 + q * kCounterSlots]);
    if (skip_empty) dm = min(dm, rmin);
    bound = (int32_t)min<int64_t>(((int64_t)dm / a.delta + 1) * a.delta, kInf);
  }
  // This workgroup's word of the next collect's scan (the previous collect,
  // which used it, has finished).
  if (threadIdx.x == 0) a.status[(par ^ 1) * G + b] = 0ull;
  __shared__ int32_t s_scan[2][kCollectBlock / kWave + 1];
  const int64_t t0 = b * a.tiles_per_block;
  const int64_t t1 = min<int64_t>(t0 + a.tiles_per_block, a.tiles);
  unsigned long long left = 0;
  int32_t dmin = INT_MAX;
  int32_t cnt = 0, dsum = 0;
  uint4 d4 = make_uint4(0u, 0u, 0u, 0u);
  uint32_t sel = 0;
  int32_t deg[kCollectPer];
  for (int64_t t = t0; t < t1; t++) {
    const int64_t v0 = t * kCollectStep + (int64_t)threadIdx.x * kCollectPer;
    sel = collect_select<VEC, true>(a, v0, bound, d4, deg, left, dmin);
    cnt += __popc(sel);
#pragma unroll
    for (int j = 0; j < kCollectPer; j++) dsum += deg[j];
  }
  int32_t tot_cnt, tot_dsum;
  int32_t pos = block_exclusive_scan<kCollectBlock>(cnt, s_scan[0], &tot_cnt);
  int32_t eoff = block_exclusive_scan<kCollectBlock>(dsum, s_scan[1], &tot_dsum);
  const uint64_t agg = ((uint64_t)tot_cnt << kPackShift) | (uint64_t)tot_dsum;
  // A workgroup with nothing selected needs no prefix, except the last one,
  // whose prefix plus its own aggregate is the frontier's size and degree sum.
  const uint64_t excl = block_exclusive_across<kCollectBlock>(a.status + par * G, (int)b, agg, tot_cnt > 0 || last);
  pos += (int32_t)(excl >> kPackShift);
  eoff += (int32_t)(excl & kPackMask);
  if (t1 - t0 == 1) {
    collect_write(a, t0 * kCollectStep + (int64_t)threadIdx.x * kCollectPer, d4, sel, deg, pos, eoff);
  } else {
    for (int64_t t = t0; t < t1; t++) {
      const int64_t v0 = t * kCollectStep + (int64_t)threadIdx.x * kCollectPer;
      sel = collect_select<VEC, false>(a, v0, bound, d4, deg, left, dmin);
      collect_write(a, v0, d4, sel, deg, pos, eoff);
    }
  }
  // What is left above the bound, for a later bucket move.
  block_counter_add<kCollectBlock>(&a.ctr[kLeftSlot + 2 * par], left);
  {
    __shared__ int32_t s_min[kCollectBlock / kWave];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) dmin = min(dmin, __shfl_xor(dmin, off));
    if (lane_id() == 0) s_min[threadIdx.x / kWave] = dmin;
    __syncthreads();
    if (threadIdx.x == 0) {
      int32_t mn = INT_MAX;
      for (int w = 0; w < kCollectBlock / kWave; w++) mn = min(mn, s_min[w]);
      if (mn != INT_MAX) atomicMin(stripe_of(&a.ctr[kLeftSlot + 2 * par + 1]), (unsigned long long)mn);
    }
  }
  if (last) {
    // Every workgroup has read the state and the previous pair of counters
    // (it published its aggregate after): clear that pair, then decide.
    if (threadIdx.x < kStripes * 2) {
      const int q = threadIdx.x >> 1, k = threadIdx.x & 1;
      a.ctr[q * kCounterSlots + kLeftSlot + 2 * (par ^ 1) + k] = k ? (unsigned long long)INT_MAX : 0ull;
    } else if (threadIdx.x < kStripes * 3) {  // the next push's smallest distance set
      a.ctr[(threadIdx.x - kStripes * 2) * kCounterSlots + kRelaxMinSlot] = (unsigned long long)INT_MAX;
    }
    if (threadIdx.x == 0) {
      const uint64_t incl = excl + agg;
      sssp_decide(a, bound, (int32_t)(incl >> kPackShift), (int64_t)(incl & kPackMask));
    }
  }
}

// The rows a pull visits, in one launch before it (gated like the pull):
// rows[v] = number of vertices u < v with dist[u] > ctl->settled, whose
// in-rows can still improve; rows[n] = their total.  The collect's shape
// (tiles of kCollectStep vertices, 16 per thread, at most kCollectGridMax
// workgroups over runs of consecutive tiles) and a single-pass scan across
// the workgroups on status words of the pull's parity (ctl->pulls, counted
// when the pull was decided).  The pull skips a merge-path tile whose rows
// are all settled: rows[i1] == rows[i0 - 1].
// VEC: dist is 16-byte aligned, so a thread's 16 distances are four 16-byte
// loads (and its 16 prefix entries four 16-byte stores) when all are in range.
template <bool VEC>
__device__ __forceinline__ uint32_t pull_row_bits(const int32_t *__restrict__ dist, int32_t n, int64_t v0,
                                                  int32_t lim) {
  uint32_t bits = 0;
  if (VEC && v0 + kCollectPer <= n) {
    const int4 *d4 = reinterpret_cast<const int4 *>(dist + v0);
#pragma unroll
    for (int q = 0; q < kCollectPer / 4; q++) {
      const int4 d = d4[q];
      bits |= (uint32_t)(d.x > lim) << (4 * q) | (uint32_t)(d.y > lim) << (4 * q + 1) |
              (uint32_t)(d.z > lim) << (4 * q + 2) | (uint32_t)(d.w > lim) << (4 * q + 3);
    }
    return bits;
  }
#pragma unroll
  for (int j = 0; j < kCollectPer; j++) {
    const int64_t v = v0 + j;
    if (v < n && dist[v] > lim) bits |= 1u << j;
  }
  return bits;
}

template <bool VEC>
__global__ void __launch_bounds__(kCollectBlock)
k_pull_rows(const SsspCtl *__restrict__ ctl, const int32_t *__restrict__ dist, int32_t n, int32_t tiles,
            int32_t tiles_per_block, int32_t *__restrict__ rows, uint64_t *__restrict__ status) {
  if (ctl->done || ctl->pull_skip) return;
  const int32_t lim = ctl->settled;
  const int par = ctl->pulls & 1;
  const int G = gridDim.x, b = blockIdx.x;
  if (threadIdx.x == 0) status[(par ^ 1) * G + b] = 0ull;  // this workgroup's word for the next pull
  __shared__ int32_t s_scan[kCollectBlock / kWave + 1];
  const int64_t t0 = (int64_t)b * tiles_per_block;
  const int64_t t1 = min<int64_t>(t0 + tiles_per_block, tiles);
  uint32_t bits = 0;
  int32_t sum = 0;
  for (int64_t t = t0; t < t1; t++) {
    bits = pull_row_bits<VEC>(dist, n, t * kCollectStep + (int64_t)threadIdx.x * kCollectPer, lim);
    sum += __popc(bits);
  }
  int32_t total;
  (void)block_exclusive_scan<kCollectBlock>(sum, s_scan, &total);
  int32_t carry = (int32_t)block_exclusive_across<kCollectBlock>(status + par * G, b, (uint64_t)total, true);
  for (int64_t t = t0; t < t1; t++) {
    const int64_t v0 = t * kCollectStep + (int64_t)threadIdx.x * kCollectPer;
    if (t1 - t0 > 1) bits = pull_row_bits<VEC>(dist, n, v0, lim);
    int32_t tile_total;
    const int32_t p = carry + block_exclusive_scan<kCollectBlock>(__popc(bits), s_scan, &tile_total);
    if (VEC && v0 + kCollectPer <= n) {
      int4 *r4 = reinterpret_cast<int4 *>(rows + v0);
#pragma unroll
      for (int q = 0; q < kCollectPer / 4; q++) {
        const int j = 4 * q;
        r4[q] = make_int4(p + __popc(bits & ((1u << j) - 1u)), p + __popc(bits & ((1u << (j + 1)) - 1u)),
                          p + __popc(bits & ((1u << (j + 2)) - 1u)), p + __popc(bits & ((1u << (j + 3)) - 1u)));
      }
    } else {
#pragma unroll
      for (int j = 0; j < kCollectPer; j++)
        if (v0 + j < n) rows[v0 + j] = p + __popc(bits & ((1u << j) - 1u));
    }
    carry += tile_total;
  }
  if (b == G - 1 && threadIdx.x == 0) rows[n] = carry;
}

// The start of a run in one launch: every element gets its final initial
// value (no fill then patch, so no ordering between threads is needed):
// dist = kInf except dist[src] = 0; the dirty bytes (16 per thread, the array
// padded to n16 x 16 bytes) clear except src's; the counters zero except the
// collect's minimum slots, at their identity in every stripe; both sets of
// scan words clear; and the run state.
__global__ void k_sssp_start(int32_t *dist, int32_t n, uint4 *dirty16, int64_t n16, unsigned long long *ctr,
                             uint64_t *status, int64_t status_words, int32_t src, SsspCtl *ctl, int32_t bound,
                             double bytes0) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  fill_except_one(dist, n, kInf, src, 0, t0, stride);
  const int64_t src16 = src / kCollectPer;
  const int sb = src % kCollectPer;  // byte of src within its 16
  const uint32_t bit = 1u << (8 * (sb & 3));
  for (int64_t i = t0; i < n16; i += stride) {
    const bool here = i == src16;
    dirty16[i] = make_uint4(here && (sb >> 2) == 0 ? bit : 0u, here && (sb >> 2) == 1 ? bit : 0u,
                            here && (sb >> 2) == 2 ? bit : 0u, here && (sb >> 2) == 3 ? bit : 0u);
  }
  for (int64_t i = t0; i < kCounterWords; i += stride) {
    const int slot = (int)(i % kCounterSlots);
    ctr[i] = slot == kLeftSlot + 1 || slot == kLeftSlot + 3 || slot == kRelaxMinSlot ? (unsigned long long)INT_MAX
                                                                                       : 0ull;
  }
  for (int64_t i = t0; i < status_words; i += stride) status[i] = 0ull;
  if (t0 == 0) {
    ctl->bound = bound;
    ctl->push_size = 0;
    ctl->tier_size = 0;
    ctl->pull_skip = 1;
    ctl->done = 0;
    ctl->advance = 0;
    ctl->epoch = 0;
    ctl->iters = 0;
    ctl->pulls = 0;
    ctl->steps = 0;
    ctl->last_nf = 0;
    ctl->prev_pull = 0;
    ctl->settled = -1;
    ctl->stall = 0;
    ctl->last_deg = 0;
    ctl->edges = 0;
    ctl->bytes = bytes0;
  }
}

}  // namespace

int sssp_near_far(gg_graph *g, int32_t src, int32_t *dist_out, uint32_t flags, gg_stats *stats) {
  GG_TRY(activate(g));
  const int32_t n = g->n;
  const int64_t m = g->m;
  hipStream_t s = g->stream;
  GG_CHECK_ARG(m < ((int64_t)1 << kPackShift), "SSSP needs fewer than 2^31 edges (m = %lld)", (long long)m);
  LbsBuffers lb;
  GG_TRY(lbs_buffers(g, &lb));
  const int64_t n_dirty = ((int64_t)n + kCollectPer - 1) / kCollectPer * kCollectPer;  // zero padding for the collect
  GG_TRY(g->vtx[3].ensure((size_t)n_dirty));
  const int32_t *pparts = nullptr;
  const void *wcv = nullptr;
  int64_t pull_alpha = g->opt.pull_alpha;
  if (pull_alpha > 0) {
    GG_TRY(pull_partition(g, &pparts));
    if (csc_values(g, &wcv) != GG_OK) {  // in-edge weights unknown: push only
      clear_error();
      pull_alpha = 0;
    }
  }
  const uint32_t *pk = nullptr;
  int32_t pk_bits = 0;
  if (pull_alpha > 0) GG_TRY(packed_in_edges(g, static_cast<const int32_t *>(wcv), &pk, &pk_bits));
  // A pull: k_pull_rows reads every distance and writes the row prefix
  // (8 n); the pull itself is counted as it runs: 8 bytes per in-edge and
  // per row of the tiles it visits (source and weight; offset and
  // distance) and 4 per distance gather (counter slots kStreamedEdgesSlot,
  // kStreamedRowsSlot, kGatheredSlot).
  const double pull_bytes = 8.0 * n;
  int32_t *dist;
  if (flags & GG_OUT_DEVICE) {
    GG_CHECK_ARG(dist_out, "dist is NULL");
    dist = dist_out;
  } else {
    GG_TRY(g->vtx[2].ensure(4 * (size_t)n));
    dist = g->vtx[2].as<int32_t>();
  }
  uint8_t *dirty = g->vtx[3].as<uint8_t>();
  int32_t *queue = g->frontier[0].as<int32_t>();
  unsigned long long *ctr = g->counters.as<unsigned long long>();
  int32_t *qsize = reinterpret_cast<int32_t *>(ctr + 8);
  const int64_t delta = g->opt.sssp_delta > 0 ? g->opt.sssp_delta : 128;
  g->prof.enabled = (flags & GG_INSTRUMENT) != 0;

  // The collect's shape: tiles of kCollectStep vertices, at most
  // kCollectGridMax workgroups, each over a run of consecutive tiles.
  const int32_t tiles = (int32_t)std::max<int64_t>(1, ((int64_t)n + kCollectStep - 1) / kCollectStep);
  const int32_t per_block = (tiles + kCollectGridMax - 1) / kCollectGridMax;
  const int32_t collect_grid = (tiles + per_block - 1) / per_block;
  // Scan words: two sets for the collect, two for k_pull_rows.
  GG_TRY(g->scan_words.ensure(4 * sizeof(uint64_t) * (size_t)collect_grid));
  GG_TRY(g->pull_rows.ensure(sizeof(int32_t) * ((size_t)n + 1)));
  GG_TRY(g->loop_ctl.ensure(sizeof(SsspCtl)));
  SsspCtl *ctl = g->loop_ctl.as<SsspCtl>();

  hipEvent_t e0, e1;
  GG_HIP(g->sssp_ev.get(&e0, &e1));
  GG_HIP(hipEventRecord(e0, s));
  {
    TimedRegion t(g->prof, "sssp.init");
    const int64_t n16 = n_dirty / kCollectPer;
    hipLaunchKernelGGL(k_sssp_start, dim3(grid_for(std::max<int64_t>(n, kCounterWords), 256, 8192)), dim3(256), 0, s,
                       dist, n, reinterpret_cast<uint4 *>(dirty), n16, ctr, g->scan_words.as<uint64_t>(),
                       4 * (int64_t)collect_grid, src, ctl, (int32_t)std::min<int64_t>(delta, kInf), 5.0 * n);
    GG_HIP(hipGetLastError());
  }
  g->prof.add_bytes("sssp.init", 5.0 * n);
  // The collect reads dist and the offsets as 16-byte vectors when both arrays allow.
  const bool collect_vec = ((reinterpret_cast<uintptr_t>(dist) | reinterpret_cast<uintptr_t>(g->dg.row_offsets)) & 15) == 0;
  const auto collect = collect_vec ? k_sssp_collect<true> : k_sssp_collect<false>;
  CollectArgs ca{dirty,       dist,        g->dg.row_offsets, n,        tiles,       per_block,
                 g->opt.iter_num, queue,   lb.scan_local,     lb.partials, lb.info, ctr,
                 lb.edge_counter, g->scan_words.as<uint64_t>(), ctl, m,      delta,       pull_alpha,
                 pull_bytes,  kSsspMaskAll};
  // mask: the iteration kinds the launches after this collect cover.
  auto launch_collect = [&](int32_t mask = kSsspMaskAll) {
    CollectArgs a = ca;
    a.mask = mask;
    hipLaunchKernelGGL(collect, dim3(collect_grid), dim3(kCollectBlock), 0, s, a);
  };
  int32_t *pull_rows = g->pull_rows.as<int32_t>();
  PullOp pull_op{g->dg.row_indices, static_cast<const int32_t *>(wcv), dist, dirty, ctr, {}, &ctl->settled,
                 pull_rows};
  PackedPullOp pk_op{pk,    pk_bits, pk_bits ? (uint32_t)((1ull << pk_bits) - 1) : 0u, dist, dirty, ctr, {},
                     &ctl->settled, pull_rows};
  // The rows a pull visits (gated like the pull).
  const bool rows_vec = ((reinterpret_cast<uintptr_t>(dist) | reinterpret_cast<uintptr_t>(pull_rows)) & 15) == 0;
  auto launch_pull_rows = [&]() {
    hipLaunchKernelGGL(rows_vec ? k_pull_rows<true> : k_pull_rows<false>, dim3(collect_grid), dim3(kCollectBlock), 0,
                       s, ctl, dist, n, tiles, per_block, pull_rows,
                       g->scan_words.as<uint64_t>() + 2 * (int64_t)collect_grid);
  };
  // One pull launch (packed in-edges when the graph has them).
  auto launch_pull = [&](const int32_t *skip) {
    if (pk)
      hipLaunchKernelGGL((k_pull_staged<PackedPullOp, int32_t>), dim3(2048), dim3(kExpandBlock), 0, s, pk_op,
                         g->dg.column_offsets, n, m, pparts, skip);
    else
      hipLaunchKernelGGL((k_pull_staged<PullOp, int32_t>), dim3(2048), dim3(kExpandBlock), 0, s, pull_op,
                         g->dg.column_offsets, n, m, pparts, skip);
  };
  const int32_t *w_out = static_cast<const int32_t *>(g->dg.edge_values);
  PushOp push_op{g->dg.row_offsets, g->dg.column_indices, w_out, nullptr, 0, 0u, dist, dirty, ctr, {}};
  // The push reads the packed words too when they are the CSR's own slots.
  const bool push_packed = pk && g->dg.row_indices == g->dg.column_indices && wcv == g->dg.edge_values;
  PackedPushOp push_pk{g->dg.row_offsets, g->dg.column_indices, w_out, pk, pk_bits,
                       pk_bits ? (uint32_t)((1ull << pk_bits) - 1) : 0u, dist, dirty, ctr, {}};
  // The LBS push over the collect's queue and offsets: partition + expand
  // (no count or scan launches: the collect wrote the offsets).
  auto launch_push = [&](auto &&timer) {
    const int32_t *fsize = &ctl->push_size;
    auto go = [&](const auto &op) {
      if (lb.vt == 2)
        lbs_partition_expand<2>(op, queue, fsize, (int64_t)n + m, lb, g->frontier[1].as<int32_t>(), qsize + 1, s,
                                timer);
      else if (lb.vt == 8)
        lbs_partition_expand<8>(op, queue, fsize, (int64_t)n + m, lb, g->frontier[1].as<int32_t>(), qsize + 1, s,
                                timer);
      else
        lbs_partition_expand<4>(op, queue, fsize, (int64_t)n + m, lb, g->frontier[1].as<int32_t>(), qsize + 1, s,
                                timer);
    };
    if (push_packed)
      go(push_pk);
    else
      go(push_op);
  };
  // The small-frontier push: one tiered launch (at most kTierItems items).
  const int tier_grid = grid_for(kTierItems, kExpandBlock);
  auto launch_tier = [&]() {
    const int32_t *fsize = &ctl->tier_size;
    if (push_packed)
      hipLaunchKernelGGL((k_tiered_expand<PackedPushOp, false>), dim3(tier_grid), dim3(kExpandBlock), 0, s, push_pk,
                         queue, fsize, g->frontier[1].as<int32_t>(), qsize + 1, g->opt.wave_degree,
                         g->opt.block_degree, lb.edge_counter);
    else
      hipLaunchKernelGGL((k_tiered_expand<PushOp, false>), dim3(tier_grid), dim3(kExpandBlock), 0, s, push_op, queue,
                         fsize, g->frontier[1].as<int32_t>(), qsize + 1, g->opt.wave_degree, g->opt.block_degree,
                         lb.edge_counter);
  };
  SsspCtl h{};
  unsigned long long gathered_seen = 0;  // host-driven runs: pull gathers counted so far
  unsigned long long streamed_seen[2] = {0, 0};  // and the pulls' visited edges and rows
  if (!g->prof.enabled) {
    // Device-driven: a replay of a captured launch sequence (per iteration:
    // the collect, which also takes the iteration's decisions; the LBS
    // partition and expand over ctl->push_size; the tiered push over
    // ctl->tier_size; the pull unless ctl->pull_skip); the host reads the
    // run state once per replay.  When the previous run on this graph took
    // at most kSpecMax collects, the first replay is that run's schedule:
    // each collect followed only by the kernels of the kind it took then.  A
    // collect that decides another kind stalls the replay (its kernels held
    // back, later collects idle); k_sssp_resume and that iteration's kernels
    // then run ahead of the general loop.  Per-kernel events (GG_INSTRUMENT)
    // need the host loop below.
    const std::vector<uint64_t> key = {
        (uint64_t)dist, (uint64_t)dirty, (uint64_t)queue, (uint64_t)ctr, (uint64_t)ctl, (uint64_t)wcv,
        (uint64_t)pparts, (uint64_t)lb.scan_local, (uint64_t)lb.partials, (uint64_t)lb.info, (uint64_t)lb.parts,
        (uint64_t)lb.edge_counter, (uint64_t)lb.expand_grid, (uint64_t)lb.vt, (uint64_t)delta,
        (uint64_t)pull_alpha, (uint64_t)g->opt.iter_num, (uint64_t)g->frontier[1].p, (uint64_t)pk,
        (uint64_t)pk_bits, (uint64_t)g->opt.wave_degree, (uint64_t)g->opt.block_degree,
        (uint64_t)g->scan_words.p, (uint64_t)collect_grid, (uint64_t)pull_rows};
    // The kernels of one iteration of the given kinds (mask), after its collect.
    auto launch_kinds = [&](int32_t mask) {
      if ((mask >> kSsspLbs) & 1) launch_push([](int) {});
      if ((mask >> kSsspTier) & 1) launch_tier();
      if (pull_alpha > 0 && ((mask >> kSsspPull) & 1)) {
        launch_pull_rows();
        launch_pull((const int32_t *)&ctl->pull_skip);
      }
    };
    // Captures what `body` launches into L (keyed by the kernel arguments).
    auto capture = [&](CapturedLoop &L, auto &&body) -> int {
      L.reset();
      hipGraph_t graph = nullptr;
      GG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      body();
      const hipError_t le = hipGetLastError();
      const hipError_t ce = hipStreamEndCapture(s, &graph);
      hipError_t ie = hipErrorUnknown;
      if (le == hipSuccess && ce == hipSuccess) ie = hipGraphInstantiate(&L.exec, graph, nullptr, nullptr, 0);
      if (graph) (void)hipGraphDestroy(graph);
      if (ie != hipSuccess) {
        L.exec = nullptr;
        return set_error(GG_ERR_HIP, "SSSP loop capture failed: %s / %s / %s", hipGetErrorString(le),
                         hipGetErrorString(ce), hipGetErrorString(ie));
      }
      L.key = key;
      return GG_OK;
    };
    // The general replay of `iters` iterations, captured on first use for
    // this graph and these buffers / options.
    auto loop_exec = [&](int iters, hipGraphExec_t *out) -> int {
      CapturedLoop &L = g->sssp_loops[iters];
      if (!L.exec || L.key != key)
        GG_TRY(capture(L, [&]() {
          for (int it = 0; it < iters; it++) {
            launch_collect();
            launch_kinds(kSsspMaskAll);
          }
        }));
      *out = L.exec;
      return GG_OK;
    };
    // The replay of the last run's schedule.
    auto scheduled = [&](hipGraphExec_t *out) -> int {
      for (SpecLoop &sp : g->sssp_spec)
        if (sp.loop.exec && sp.loop.key == key && sp.sched == g->sssp_sched) {
          *out = sp.loop.exec;
          return GG_OK;
        }
      SpecLoop &sp = g->sssp_spec[g->sssp_spec_next];
      g->sssp_spec_next = (g->sssp_spec_next + 1) % kSpecCache;
      sp.sched.clear();
      GG_TRY(capture(sp.loop, [&]() {
        for (uint8_t k : g->sssp_sched) {
          const int32_t mask = k == kSsspCollectOnly ? 0 : 1 << k;
          launch_collect(mask);
          launch_kinds(mask);
        }
      }));
      sp.sched = g->sssp_sched;
      *out = sp.loop.exec;
      return GG_OK;
    };
    // The held-back iteration of a stalled replay.
    auto resume_exec = [&](hipGraphExec_t *out) -> int {
      CapturedLoop &L = g->sssp_resume;
      if (!L.exec || L.key != key)
        GG_TRY(capture(L, [&]() {
          hipLaunchKernelGGL(k_sssp_resume, dim3(1), dim3(64), 0, s, ctl, lb.info);
          launch_kinds(kSsspMaskAll);
        }));
      *out = L.exec;
      return GG_OK;
    };
    // The schedule is replayed once two runs in a row took it (repeated
    // queries); while schedules vary from query to query (sources in turn),
    // a stall would cost a second synchronisation, so the first replay is
    // then the general loop, as long as the previous run.  Later replays run
    // kLoopIters.
    const bool use_sched = !g->sssp_sched.empty() && g->sssp_sched_repeated;
    int iters = g->sssp_steps_hint > 0 ? std::min(g->sssp_steps_hint, kSsspReplayMax) : kLoopIters;
    // After each replay, one synchronisation reads the run state and (when
    // asked for) the reach statistics, enqueued on the chance that the replay
    // ends the run (otherwise they are dropped).
    for (int replay = 0;; replay++) {
      if (replay == 0 && use_sched) {
        hipGraphExec_t exec = nullptr;
        GG_TRY(scheduled(&exec));
        GG_HIP(hipGraphLaunch(exec, s));
      } else {
        if (h.stall) {
          hipGraphExec_t rexec = nullptr;
          GG_TRY(resume_exec(&rexec));
          GG_HIP(hipGraphLaunch(rexec, s));
        }
        hipGraphExec_t exec = nullptr;
        GG_TRY(loop_exec(iters, &exec));
        GG_HIP(hipGraphLaunch(exec, s));
        iters = kLoopIters;
      }
      GG_HIP(hipEventRecord(e1, s));
      if (stats) GG_TRY(reach_launch(g, dist, kInf));
      GG_TRY(enqueue_handoff(g, dist));  // the caller's result, from dist itself
      GG_TRY(counters_to_host_async(g, 34, 1ull << 17, ctl, sizeof(SsspCtl)));
      GG_HIP(hipStreamSynchronize(s));
      std::memcpy(&h, g->h64() + kCounterSlots, sizeof(h));
      if (replay == 0 && use_sched) {
        // The scheduled replay ended early (a stall) or ran short: the next
        // one runs the collects the previous run still had left.
        const int left = g->sssp_steps_hint - h.steps;
        iters = left > 0 ? std::min(left, kSsspReplayMax) : kLoopIters;
      }
      if (h.done) {
        g->sssp_steps_hint = h.steps;
        const std::vector<uint8_t> ran(h.kinds, h.kinds + std::min(h.steps, kSpecMax));
        g->sssp_sched_repeated = h.steps <= kSpecMax && ran == g->sssp_sched;
        if (h.steps <= kSpecMax)
          g->sssp_sched = ran;
        else
          g->sssp_sched.clear();
        g->handoff.done = g->handoff.armed;
        break;
      }
    }
  } else {
    // Host-driven, one kernel event each: the collect (with the same
    // on-device decisions), a read-back of the run state, then the launch it
    // asks for.
    auto timer = [&](int) {
      g->prof.end("sssp.lbs_partition");
      g->prof.begin("sssp.push");
    };
    for (;;) {
      {
        TimedRegion t(g->prof, "sssp.collect");
        launch_collect();
        GG_HIP(hipGetLastError());
      }
      GG_TRY(counters_to_host_async(g, 0, 0, ctl, sizeof(SsspCtl)));
      GG_HIP(hipStreamSynchronize(s));
      std::memcpy(&h, g->h64() + kCounterSlots, sizeof(h));
      if (h.done) break;
      g->prof.add_launch_bytes("sssp.collect", n + 16.0 * h.last_nf);
      const double pb = 16.0 * h.last_nf + 12.0 * (double)h.last_deg;
      if (!h.pull_skip) {
        {
          TimedRegion t(g->prof, "sssp.pull_rows");
          launch_pull_rows();
          GG_HIP(hipGetLastError());
        }
        g->prof.add_launch_bytes("sssp.pull_rows", pull_bytes);
        {
          TimedRegion t(g->prof, "sssp.pull");
          launch_pull((const int32_t *)nullptr);
          GG_HIP(hipGetLastError());
        }
        // This launch's share of the pull counters (they accumulate over the run).
        GG_TRY(read_counters(g, kStreamedRowsSlot + 1));
        const unsigned long long gsum = g->h64()[kGatheredSlot], esum = g->h64()[kStreamedEdgesSlot],
                                 rsum = g->h64()[kStreamedRowsSlot];
        g->prof.add_launch_bytes("sssp.pull", 4.0 * (double)(gsum - gathered_seen) +
                                                  8.0 * (double)(esum - streamed_seen[0]) +
                                                  8.0 * (double)(rsum - streamed_seen[1]));
        gathered_seen = gsum;
        streamed_seen[0] = esum;
        streamed_seen[1] = rsum;
      } else if (h.tier_size > 0) {
        {
          TimedRegion t(g->prof, "sssp.push_small");
          launch_tier();
          GG_HIP(hipGetLastError());
        }
        g->prof.add_launch_bytes("sssp.push_small", pb);
      } else if (h.push_size > 0) {
        g->prof.begin("sssp.lbs_partition");
        launch_push(timer);
        g->prof.end("sssp.push");
        GG_HIP(hipGetLastError());
        g->prof.add_launch_bytes("sssp.push", pb);
      }
    }
    GG_HIP(hipEventRecord(e1, s));
    GG_TRY(read_counters(g, 8));
  }
  const int64_t improved = g->h64()[4];
  const int64_t gathered = g->h64()[kGatheredSlot];
  const int64_t pulled_edges = g->h64()[kStreamedEdgesSlot], pulled_rows = g->h64()[kStreamedRowsSlot];
  // dist + dirty writes of every improvement; the pulls' visited tiles and gathers
  const double bytes = h.bytes + 8.0 * improved + 4.0 * gathered + 8.0 * pulled_edges + 8.0 * pulled_rows;
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  finish_profile(g);
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    if (!g->prof.enabled) {
      stats->reached_vertices = g->h64()[32];
      stats->reached_edges = g->h64()[33];
    } else {
      GG_TRY(reach_stats(g, dist, kInf, stats));
    }
    stats->iterations = h.iters;
    stats->pull_iterations = h.pulls;
    stats->edges_inspected = h.edges + pulled_edges;
    stats->alg_bytes = bytes;
    stats->gpu_ms = ms;
  }
  if (flags & GG_OUT_DEVICE) return GG_OK;
  if (g->handoff.done) return GG_OK;  // the caller's buffer was written with the last read-back
  return copy_out(g, dist_out, dist, 4 * (size_t)n, flags);
}

}  // namespace gg
