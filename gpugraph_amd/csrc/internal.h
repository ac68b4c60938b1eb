// Internal definition of the gg_graph handle and helpers shared by the
// library's HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "common.h"
#include "gasengine/enactor_vertex_centric.h"

namespace gg {

#define GG_HIP(expr)                                                                              \
  do {                                                                                            \
    hipError_t _e = (expr);                                                                       \
    if (_e != hipSuccess)                                                                         \
      return ::gg::set_error(_e == hipErrorOutOfMemory ? GG_ERR_OOM : GG_ERR_HIP, "%s failed: %s (%s:%d)", \
                             #expr, hipGetErrorString(_e), __FILE__, __LINE__);                   \
  } while (0)

#define GG_TRY(expr)               \
  do {                             \
    int _rc = (expr);              \
    if (_rc != GG_OK) return _rc;  \
  } while (0)

// Device buffer owned by a graph handle.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  // Grow to at least `want` bytes (contents not preserved).
  int ensure(size_t want) {
    if (want == 0) want = 4;
    if (bytes >= want) return GG_OK;
    reset();
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
      p = nullptr;
      return set_error(GG_ERR_OOM, "hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
    }
    bytes = want;
    return GG_OK;
  }
  template <class T>
  T *as() const {
    return reinterpret_cast<T *>(p);
  }
  void swap(DevBuf &o) {
    std::swap(p, o.p);
    std::swap(bytes, o.bytes);
  }
};

// HIP-event timer per kernel class (GG_INSTRUMENT).
class EventProfiler : public GASengine::KernelTimer {
 public:
  struct Entry {
    int64_t launches = 0;
    double ms = 0.0;
    double bytes = 0.0;
  };
  hipStream_t stream = nullptr;
  bool enabled = false;
  std::map<std::string, Entry> table;
  // Per-launch record (resolved launches, oldest first, capped at kMaxLaunches).
  struct Launch {
    std::string name;
    double ms, bytes;
  };
  static constexpr size_t kMaxLaunches = 1 << 16;
  std::vector<Launch> launches;

  ~EventProfiler() override;
  void begin(const char *name) override;
  void end(const char *name) override;
  void add_bytes(const char *name, double bytes) { if (enabled) table[name].bytes += bytes; }
  // Same, and attribute the bytes to the latest launch of that class.
  void add_launch_bytes(const char *name, double bytes);
  // Resolve recorded event pairs (stream must be synchronised).
  void flush();
  void reset();

 private:
  struct Pending {
    std::string name;
    hipEvent_t a, b;
    double bytes;
  };
  std::vector<Pending> pending_;
  std::vector<hipEvent_t> pool_;
  hipEvent_t open_ = nullptr;
  hipEvent_t take();
};

// An instantiated capture of a device-driven iteration loop, cached on a
// graph handle with the kernel arguments it was captured with (`key`); a
// call whose arguments differ captures again.
// Longest device-driven SSSP replay that is captured (iterations); the graph
// keeps one captured replay per length 1..kSsspReplayMax.
constexpr int kSsspReplayMax = 16;

struct CapturedLoop {
  hipGraphExec_t exec = nullptr;
  std::vector<uint64_t> key;
  CapturedLoop() = default;
  CapturedLoop(const CapturedLoop &) = delete;
  CapturedLoop &operator=(const CapturedLoop &) = delete;
  ~CapturedLoop() { reset(); }
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
    key.clear();
  }
};

// A device-driven loop captured for one schedule: the kinds of the
// iterations (levels) a previous run on the same graph took, each slot
// holding only the kernels its kind launches.  The loop's controller
// kernels know each slot's kind; an iteration that needs another kind
// stalls the replay (nothing more runs in it) and the host continues with
// the general loop, whose slots hold every kernel.
constexpr int kSpecMax = 32;    // longest schedule recorded (iterations)
constexpr int kSpecCache = 4;   // schedules kept captured per graph and algorithm
struct SpecLoop {
  CapturedLoop loop;
  std::vector<uint8_t> sched;
};

// Two events timing a run, created on first use and kept with the graph
// (creating and destroying them on every call costs host time per query).
struct RunEvents {
  hipEvent_t e[2] = {nullptr, nullptr};
  RunEvents() = default;
  RunEvents(const RunEvents &) = delete;
  RunEvents &operator=(const RunEvents &) = delete;
  ~RunEvents() {
    for (hipEvent_t x : e)
      if (x) (void)hipEventDestroy(x);
  }
  hipError_t get(hipEvent_t *e0, hipEvent_t *e1) {
    for (hipEvent_t &x : e)
      if (!x) {
        const hipError_t r = hipEventCreate(&x);
        if (r != hipSuccess) return r;
      }
    *e0 = e[0];
    *e1 = e[1];
    return hipSuccess;
  }
};

// Two events owned by one call: destroyed on every return path.
struct ScopedEvents {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  ScopedEvents() = default;
  ScopedEvents(const ScopedEvents &) = delete;
  ScopedEvents &operator=(const ScopedEvents &) = delete;
  ~ScopedEvents() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
  hipError_t create() {
    hipError_t r = hipEventCreate(&e0);
    return r != hipSuccess ? r : hipEventCreate(&e1);
  }
};

// A captured graph and its executable owned by one call.
struct ScopedGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ScopedGraph() = default;
  ScopedGraph(const ScopedGraph &) = delete;
  ScopedGraph &operator=(const ScopedGraph &) = delete;
  ~ScopedGraph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

// Scoped helper for kernels launched outside the generic enactor.
struct TimedRegion {
  EventProfiler &p;
  const char *name;
  TimedRegion(EventProfiler &p_, const char *n) : p(p_), name(n) { p.begin(n); }
  ~TimedRegion() { p.end(name); }
};

// Phase callback for GASengine::lbs_launch: times the four LBS kernels as
// "<prefix>.lbs_count", ".lbs_scan", ".lbs_partition", "<prefix>.<expand_name>".
struct LbsPhaseTimer {
  EventProfiler &p;
  std::string names[4];
  LbsPhaseTimer(EventProfiler &p_, const char *prefix, const char *expand_name) : p(p_) {
    names[0] = std::string(prefix) + ".lbs_count";
    names[1] = std::string(prefix) + ".lbs_scan";
    names[2] = std::string(prefix) + ".lbs_partition";
    names[3] = std::string(prefix) + "." + expand_name;
  }
  void operator()(int phase) {
    if (phase > 0) p.end(names[phase - 1].c_str());
    if (phase < 4) p.begin(names[phase].c_str());
  }
};

}  // namespace gg

struct gg_graph {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int32_t n = 0;
  int64_t m = 0;
  int directed = 0;
  int32_t value_dtype = GG_VALUES_NONE;
  gg_options opt{};

  // Resident topology.
  gg::DevBuf ro, col, val, co, ri, eid;
  GASengine::DeviceGraph dg;  // view of the buffers above

  // Shared scratch for the specialised paths (BFS, CC hooking, PageRank).
  gg::DevBuf frontier[2], bits[4], vtx[4], scan_local, partials, parts, info, counters;
  gg::DevBuf pr[8];    // PageRank vectors / CSC-ordered transition weights
  gg::DevBuf batch[8]; // batched-instance state
  gg::DevBuf tree[2];  // gg_bfs_tree: device depth and parent arrays
  gg::DevBuf msbfs[4]; // batched BFS: seen, frontier words (2 n), next, sizes + sources
  gg::DevBuf ms_chunks;          // batched BFS: the long in-rows cut into chunks (built once)
  int64_t ms_chunk_count = -1;   // -1: not built yet
  // Degree reordering (GG_GRAPH_REORDER): order[new] = old, new_id[old] =
  // new, emap[new slot] = old slot; results in internal order before mapping;
  // staging for host outputs; host copy of new_id for source ids.
  bool reordered = false;
  gg::DevBuf re_order, re_new, re_emap, re_out[2], re_in[2], re_stage;
  std::vector<int32_t> h_new_id;
  // Result hand-off to the caller, armed by a renumbered query: a
  // device-driven core enqueues it (enqueue_handoff) ahead of its last
  // read-back, so the mapping shares that synchronisation; `done` says the
  // caller's buffer holds the result.
  struct Handoff {
    void *dst = nullptr;
    uint32_t flags = 0;
    bool armed = false, done = false;
  } handoff;
  // Degree bins of the vertices for row reductions: [0] over CSC in-edges,
  // [1] over CSR out-edges; lists of small / mid / large rows.
  gg::DevBuf bins[2][3];
  int32_t bin_count[2][3] = {{0, 0, 0}, {0, 0, 0}};
  bool bins_ready[2] = {false, false};
  int64_t pr_weights_version = -1;  // values version the PageRank transition weights were built for
  // XCD-sliced row structure per direction ([0] CSC, [1] CSR): slice offsets
  // (8n + 1), gathered ids and original slots in slice order, long-row lists,
  // slice-ordered transition weights (F32 graphs) and the partials buffer.
  gg::DevBuf sl_off[2], sl_idx[2], sl_perm[2], sl_long[2], sl_val[2], sl_part;
  int32_t sl_mid_base[2][9] = {}, sl_large_base[2][9] = {};
  bool sl_ready[2] = {false, false};
  int64_t sl_val_version[2] = {-1, -1};
  int32_t *h_pinned = nullptr;  // pinned readback area (64 x int64 counters, then a block of read-back words)
  int64_t *h_dev = nullptr;     // the same area as kernels address it
  bool pr_cache_valid = false;

  // Graph-derived caches: the static merge-path partition of the CSC (pull
  // steps over all vertices) and the edge values permuted into CSC order
  // (directed graphs; undirected graphs read the CSR values directly).
  gg::DevBuf pull_parts, w_csc;
  bool pull_parts_ready = false;
  // Bitmap of the vertices without in-edges (BFS starts with them visited),
  // and the first four in-neighbours of every vertex (16 bytes each, -1 past
  // a shorter row) for bottom-up BFS levels.
  gg::DevBuf noin, first4;
  bool noin_ready = false, first4_ready = false;
  // The same partition over (vertices, CSR out-edges), and the per-tile
  // carries of the segmented row sums.
  gg::DevBuf csr_parts, seg_carry;
  bool csr_parts_ready = false;
  // Packed CSC in-edges for the SSSP pull (source << pk_wbits | weight),
  // built per values version when every weight and source id fits.
  gg::DevBuf pk_edges;
  int64_t pk_version = -1;
  int32_t pk_wbits = 0;  // 0: not packable
  int64_t values_version = 0, w_csc_version = -1;
  // Smallest / largest int32 edge value, for values version wr_version.
  int32_t wr[2] = {0, 0};
  int64_t wr_version = -1;
  // Undirected graphs: rows sorted and symmetry confirmed (mirror_ok); the
  // mirrored entries' values differ (values_asymmetric; mirror permutation
  // then kept in eid).
  bool mirror_ok = false, values_asymmetric = false;

  // Generic-program state kept across runs (one per program type).
  std::shared_ptr<void> prog_state[8];

  // Device-driven BFS / SSSP loops: captured launch sequences and their
  // device-side loop state.
  gg::RunEvents bfs_ev, sssp_ev;
  gg::CapturedLoop bfs_loop[2];  // general BFS loop, by the parity of its first level
  gg::CapturedLoop sssp_loops[gg::kSsspReplayMax + 1];  // by iterations per replay
  gg::CapturedLoop sssp_resume;     // the iteration kernels after a stalled SSSP collect
  int sssp_steps_hint = 0;          // loop iterations the last device-driven SSSP took
  // Schedules of the last device-driven runs and their captured loops.
  std::vector<uint8_t> bfs_sched, sssp_sched;
  gg::SpecLoop bfs_spec[gg::kSpecCache], sssp_spec[gg::kSpecCache];
  int bfs_spec_next = 0, sssp_spec_next = 0;  // slot the next capture replaces
  bool sssp_sched_repeated = false;  // the last two device-driven SSSP runs took sssp_sched
  gg::DevBuf loop_ctl;
  // Aggregate words of single-pass scans across workgroups (the SSSP
  // collect: two sets, by collect parity).
  gg::DevBuf scan_words;
  gg::DevBuf pull_rows;  // SSSP pull: prefix count of the vertices whose in-rows can still change

  gg::EventProfiler prof;

  int64_t *h64() { return reinterpret_cast<int64_t *>(h_pinned); }
};

namespace gg {

// Make the graph's device current; returns status.
int activate(gg_graph *g);
// Copy n elements of a result from device array `src` to `dst` (host or
// device per flags).
int copy_out(gg_graph *g, void *dst, const void *src, size_t bytes, uint32_t flags);
// reached vertices / reached-edge (degree) sums for a result array: counts
// entries != sentinel.  Synchronises the stream.
int reach_stats(gg_graph *g, const int32_t *result, int32_t sentinel, gg_stats *st);
// The same count enqueued only: counters 32 (vertices) and 33 (degree sum)
// hold it once the stream reaches that point.
int reach_launch(gg_graph *g, const int32_t *result, int32_t sentinel);
// Copy counters [0, count) of the graph's striped counter array
// (GASengine::kStripes stripes of kCounterSlots words) to h64()[0, count),
// each folded over its stripes: a sum, or a minimum for the bits set in
// min_mask.  Synchronises the stream.
int read_counters(gg_graph *g, int count, uint64_t min_mask = 0);
// read_counters without the synchronisation, plus `extra_bytes` (a multiple
// of 8, 8-byte aligned) copied from device memory `extra` to
// h64() + kCounterSlots: one kernel writes both into the pinned area, valid
// once the stream has passed it.
int counters_to_host_async(gg_graph *g, int count, uint64_t min_mask, const void *extra, size_t extra_bytes);
// Finish a run: resolve profiler events.
void finish_profile(gg_graph *g);
// Frontier queues, scan/partition buffers and counters of the specialised
// paths; fills *b for GASengine::lbs_launch (edge counter = counters[2]).
int lbs_buffers(gg_graph *g, GASengine::LbsBuffers *b);
// Static merge-path partition of (vertices, CSC in-edges) for k_pull_reduce.
int pull_partition(gg_graph *g, const int32_t **parts);
// The same over (vertices, CSR out-edges).
int csr_partition(gg_graph *g, const int32_t **parts);
// Edge values in CSC order (4-byte values): the CSR array itself for
// undirected graphs, a cached permuted copy for directed ones.
int csc_values(gg_graph *g, const void **vals);
// Degree reordering (reorder.hip).  build_reorder renumbers the validated
// device CSR (ro0, col0, val0) into g->ro / col / val; the map helpers move
// 4-byte vertex or edge arrays between caller order and internal order
// (host or device side per GG_OUT_DEVICE in flags).
int build_reorder(gg_graph *g, const int32_t *ro0, const int32_t *col0, const void *val0);
int32_t map_vertex(const gg_graph *g, int32_t v);
int reorder_out_vertices(gg_graph *g, const void *in, void *dst, uint32_t flags);
// Enqueues the armed hand-off of `result` (internal order) on g->stream, no
// synchronisation; a no-op when none is armed.
int enqueue_handoff(gg_graph *g, const void *result);
int reorder_in_vertices(gg_graph *g, const void *src, uint32_t flags, DevBuf &buf, const void **out);
int reorder_in_edges(gg_graph *g, const void *src, uint32_t flags, void *dst_dev);
int reorder_out_edges(gg_graph *g, const void *in, void *dst, uint32_t flags);
int reorder_out_cc(gg_graph *g, const int32_t *label, int32_t *dst, uint32_t flags);
// Smallest and largest int32 edge value (0, 0 for other value types),
// cached per values version.
int weight_range(gg_graph *g, int32_t *min_w, int32_t *max_w);
// GG_ERR_INVALID when an int32 edge weight is negative: shortest paths
// then need not exist (a negative undirected edge is a negative cycle), and
// runs that go to quiescence would not end.
int require_nonnegative_weights(gg_graph *g, const char *what);
// Batched multi-source BFS (batch_bfs.hip): num sources, 64 per pass; depths
// instance-major (num x n) to host or device memory per flags; statistics
// added to *acc.
int batch_bfs(gg_graph *g, const int32_t *sources, int32_t num, int32_t *depth_out, uint32_t flags, gg_stats *acc);
// Symmetry / value-symmetry check of an undirected graph (sets mirror_ok,
// values_asymmetric); GG_ERR_INVALID when the CSR is provably not symmetric.
int check_undirected(gg_graph *g);

}  // namespace gg
