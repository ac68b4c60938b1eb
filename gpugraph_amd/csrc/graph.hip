// Library core: errors, graph handles (device CSR + CSC), options, the
// per-kernel event profiler and result helpers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <new>
#include <vector>

#include "internal.h"

// ---------------------------------------------------------------- errors --
namespace {
thread_local char g_err[512] = "";
}

namespace gg {

int set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

void clear_error() { g_err[0] = '\0'; }

int activate(gg_graph *g) {
  int cur = -1;
  GG_HIP(hipGetDevice(&cur));
  if (cur != g->device) GG_HIP(hipSetDevice(g->device));
  return GG_OK;
}

int copy_out(gg_graph *g, void *dst, const void *src, size_t bytes, uint32_t flags) {
  if (!dst || bytes == 0) return GG_OK;
  if (flags & GG_OUT_DEVICE) {
    if (dst != src) GG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, g->stream));
    GG_HIP(hipStreamSynchronize(g->stream));
  } else {
    GG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g->stream));
    GG_HIP(hipStreamSynchronize(g->stream));
  }
  return GG_OK;
}

__global__ void k_reach(const int32_t *__restrict__ res, int32_t sentinel, const int32_t *__restrict__ ro, int32_t n,
                        unsigned long long *__restrict__ out) {
  unsigned long long cnt = 0, deg = 0;
  for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    if (res[v] != sentinel) {
      cnt++;
      deg += (unsigned long long)(ro[v + 1] - ro[v]);
    }
  GASengine::block_counter_add<256>(&out[0], cnt);
  GASengine::block_counter_add<256>(&out[1], deg);
}

int reach_launch(gg_graph *g, const int32_t *result, int32_t sentinel) {
  GG_TRY(g->counters.ensure(GASengine::kCounterWords * sizeof(unsigned long long)));
  auto *c = g->counters.as<unsigned long long>();
  GG_HIP(GASengine::slot_resets_async(c, {{32, 0ull}, {33, 0ull}}, g->stream));
  hipLaunchKernelGGL(k_reach, dim3(GASengine::grid_for(g->n, 256, 2048)), dim3(256), 0, g->stream, result, sentinel,
                     g->dg.row_offsets, g->n, c + 32);
  GG_HIP(hipGetLastError());
  return GG_OK;
}

int reach_stats(gg_graph *g, const int32_t *result, int32_t sentinel, gg_stats *st) {
  if (!st) return GG_OK;
  GG_TRY(reach_launch(g, result, sentinel));
  GG_TRY(read_counters(g, 34));
  st->reached_vertices = g->h64()[32];
  st->reached_edges = g->h64()[33];
  return GG_OK;
}

// One workgroup of kCounterSlots threads: counter k < count folded over its
// stripes into host[k] (a sum, or a minimum for the bits of min_mask), and
// `words` 8-byte words of `extra` into host[kCounterSlots + j].  `host` is
// the graph's pinned area, written by the kernel itself: no copy-engine
// transfer per read-back.
__global__ void k_counters_to_host(const unsigned long long *__restrict__ ctr, int count, uint64_t min_mask,
                                   const int64_t *__restrict__ extra, int words, int64_t *host) {
  using GASengine::kCounterSlots;
  using GASengine::kStripes;
  const int t = threadIdx.x;
  if (t < count) {
    const bool is_min = (min_mask >> t) & 1u;
    unsigned long long v = ctr[t];
    for (int q = 1; q < kStripes; q++) {
      const unsigned long long x = ctr[(size_t)q * kCounterSlots + t];
      v = is_min ? (x < v ? x : v) : v + x;
    }
    host[t] = (int64_t)v;
  }
  for (int j = t; j < words; j += blockDim.x) host[kCounterSlots + j] = extra[j];
}

int counters_to_host_async(gg_graph *g, int count, uint64_t min_mask, const void *extra, size_t extra_bytes) {
  using GASengine::kCounterSlots;
  GG_CHECK_ARG(count >= 0 && count <= kCounterSlots, "counter range");
  GG_CHECK_ARG(extra_bytes % 8 == 0 && extra_bytes / 8 <= (size_t)GASengine::kCounterWords &&
                   (extra_bytes == 0 || (extra && reinterpret_cast<uintptr_t>(extra) % 8 == 0)),
               "read-back block");
  if (count > 0) GG_TRY(g->counters.ensure(GASengine::kCounterWords * sizeof(unsigned long long)));
  hipLaunchKernelGGL(k_counters_to_host, dim3(1), dim3(kCounterSlots), 0, g->stream,
                     g->counters.as<unsigned long long>(), count, min_mask, static_cast<const int64_t *>(extra),
                     (int)(extra_bytes / 8), g->h_dev);
  GG_HIP(hipGetLastError());
  return GG_OK;
}

int read_counters(gg_graph *g, int count, uint64_t min_mask) {
  GG_TRY(counters_to_host_async(g, count, min_mask, nullptr, 0));
  GG_HIP(hipStreamSynchronize(g->stream));
  return GG_OK;
}

void finish_profile(gg_graph *g) {
  if (g->prof.enabled) g->prof.flush();
}

int pull_partition(gg_graph *g, const int32_t **parts) {
  if (!g->pull_parts_ready) {
    GG_TRY(g->pull_parts.ensure(4 * GASengine::lbs_partition_entries(g->n, g->m)));
    const int64_t tiles = ((int64_t)g->n + g->m) / GASengine::kLbsTile + 2;
    hipLaunchKernelGGL(GASengine::k_static_partition<0>, dim3(GASengine::grid_for(tiles, 256, 4096)), dim3(256), 0,
                       g->stream, g->dg.column_offsets, g->n, g->m, g->pull_parts.as<int32_t>());
    GG_HIP(hipGetLastError());
    g->pull_parts_ready = true;
  }
  *parts = g->pull_parts.as<int32_t>();
  return GG_OK;
}

int csr_partition(gg_graph *g, const int32_t **parts) {
  if (!g->csr_parts_ready) {
    GG_TRY(g->csr_parts.ensure(4 * GASengine::lbs_partition_entries(g->n, g->m)));
    const int64_t tiles = ((int64_t)g->n + g->m) / GASengine::kLbsTile + 2;
    hipLaunchKernelGGL(GASengine::k_static_partition<0>, dim3(GASengine::grid_for(tiles, 256, 4096)), dim3(256), 0,
                       g->stream, g->dg.row_offsets, g->n, g->m, g->csr_parts.as<int32_t>());
    GG_HIP(hipGetLastError());
    g->csr_parts_ready = true;
  }
  *parts = g->csr_parts.as<int32_t>();
  return GG_OK;
}

// Mirror check of an undirected CSR, run over every (v, slot) with the pull
// merge path: for slot k of v's row (neighbour u, its j-th occurrence there)
// the mirror is the j-th occurrence of v in u's row.  flags: 1 = a row is
// not sorted (no check possible), 2 = an entry has no mirror, 4 = a mirrored
// pair carries different values.
struct MirrorOp {
  const int32_t *ro;
  const int32_t *col;
  const uint32_t *val;
  int32_t *mir;
  unsigned int *flags;

  __device__ static int32_t lower_bound(const int32_t *a, int32_t lo, int32_t hi, int32_t x) {
    while (lo < hi) {
      const int32_t mid = (lo + hi) >> 1;
      if (a[mid] < x)
        lo = mid + 1;
      else
        hi = mid;
    }
    return lo;
  }
  __device__ __forceinline__ uint32_t identity() { return 0u; }
  __device__ __forceinline__ uint32_t edge(int32_t v, int32_t k) {
    const int32_t u = col[k];
    const int32_t b = ro[v];
    if (k > b && col[k - 1] > u) {
      atomicOr(flags, 1u);
      return 0u;
    }
    const int32_t j = k - lower_bound(col, b, k, u);  // occurrence index of u in v's row
    const int32_t pos = lower_bound(col, ro[u], ro[u + 1], v) + j;
    if (pos >= ro[u + 1] || col[pos] != v) {
      atomicOr(flags, 2u);
      return 0u;
    }
    mir[k] = pos;
    if (val && val[pos] != val[k]) atomicOr(flags, 4u);
    return 0u;
  }
  __device__ __forceinline__ void combine(uint32_t *, uint32_t) {}
  __device__ __forceinline__ void item_done(int32_t, uint32_t, bool) {}
  __device__ __forceinline__ void finish() {}
};

__global__ void k_permute4(const uint32_t *__restrict__ in, const int32_t *__restrict__ idx, uint32_t *__restrict__ out,
                           int64_t m) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += (int64_t)gridDim.x * blockDim.x)
    out[k] = in[idx[k]];
}

int csc_values(gg_graph *g, const void **vals) {
  const int32_t *perm = g->directed ? g->dg.csc_edge_ids : (g->values_asymmetric ? g->eid.as<int32_t>() : nullptr);
  if (!g->directed && g->values_asymmetric && !g->mirror_ok)
    return set_error(GG_ERR_UNSUPPORTED, "undirected graph with unsorted rows: in-edge values unknown");
  if (!perm) {
    *vals = g->val.p;
    return GG_OK;
  }
  if (g->w_csc_version != g->values_version) {
    GG_TRY(g->w_csc.ensure(4 * (size_t)std::max<int64_t>(g->m, 1)));
    hipLaunchKernelGGL(k_permute4, dim3(GASengine::grid_for(g->m, 256, 16384)), dim3(256), 0, g->stream,
                       g->val.as<uint32_t>(), perm, g->w_csc.as<uint32_t>(), g->m);
    GG_HIP(hipGetLastError());
    g->w_csc_version = g->values_version;
  }
  *vals = g->w_csc.p;
  return GG_OK;
}

// Smallest and largest int32 edge value: per-thread running min / max,
// wave shuffles, then one atomicMin / atomicMax per wave.
__global__ void __launch_bounds__(256) k_minmax_i32(const int32_t *__restrict__ w, int64_t m, int32_t *__restrict__ mm) {
  int32_t lo = INT_MAX, hi = INT_MIN;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += (int64_t)gridDim.x * blockDim.x) {
    const int32_t x = w[k];
    lo = min(lo, x);
    hi = max(hi, x);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    lo = min(lo, __shfl_xor(lo, off));
    hi = max(hi, __shfl_xor(hi, off));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMin(mm, lo);
    atomicMax(mm + 1, hi);
  }
}

__global__ void k_minmax_init(int32_t *mm) {
  mm[0] = INT_MAX;
  mm[1] = INT_MIN;
}

int weight_range(gg_graph *g, int32_t *min_w, int32_t *max_w) {
  if (g->value_dtype != GG_VALUES_I32 || g->m == 0) {
    *min_w = 0;
    *max_w = 0;
    return GG_OK;
  }
  if (g->wr_version != g->values_version) {
    hipStream_t s = g->stream;
    DevBuf mm;
    GG_TRY(mm.ensure(2 * sizeof(int32_t)));
    int32_t *d = mm.as<int32_t>();
    hipLaunchKernelGGL(k_minmax_init, dim3(1), dim3(1), 0, s, d);
    hipLaunchKernelGGL(k_minmax_i32, dim3(GASengine::grid_for(g->m, 256, 4096)), dim3(256), 0, s,
                       g->val.as<int32_t>(), g->m, d);
    GG_HIP(hipGetLastError());
    GG_HIP(hipMemcpyAsync(g->wr, d, sizeof(g->wr), hipMemcpyDeviceToHost, s));
    GG_HIP(hipStreamSynchronize(s));
    g->wr_version = g->values_version;
  }
  *min_w = g->wr[0];
  *max_w = g->wr[1];
  return GG_OK;
}

int require_nonnegative_weights(gg_graph *g, const char *what) {
  int32_t lo = 0, hi = 0;
  GG_TRY(weight_range(g, &lo, &hi));
  if (lo < 0)
    return set_error(GG_ERR_INVALID, "%s needs non-negative edge weights (smallest weight %d)", what, lo);
  return GG_OK;
}

// Undirected graphs: confirm the CSR is symmetric (every u->v has a v->u)
// and record whether mirrored entries carry equal values; keeps the mirror
// permutation in g->eid when they do not.  Rows that are not sorted cannot
// be checked and leave mirror_ok = false.
int check_undirected(gg_graph *g) {
  const int32_t *parts = nullptr;
  GG_TRY(pull_partition(g, &parts));
  GG_TRY(g->eid.ensure(4 * (size_t)std::max<int64_t>(g->m, 1)));
  GG_TRY(g->counters.ensure(GASengine::kCounterWords * sizeof(unsigned long long)));
  unsigned int *flags = reinterpret_cast<unsigned int *>(g->counters.as<unsigned long long>() + 40);
  GG_HIP(GASengine::fill_async<unsigned int>(flags, 0u, 1, g->stream));
  MirrorOp op{g->dg.row_offsets, g->dg.column_indices, g->val.as<uint32_t>(), g->eid.as<int32_t>(), flags};
  hipLaunchKernelGGL((GASengine::k_pull_reduce<MirrorOp, uint32_t>), dim3(2048), dim3(GASengine::kExpandBlock), 0,
                     g->stream, op, g->dg.row_offsets, g->n, g->m, parts, (const int32_t *)nullptr);
  GG_HIP(hipGetLastError());
  unsigned int h = 0;
  GG_HIP(hipMemcpyAsync(&h, flags, sizeof(h), hipMemcpyDeviceToHost, g->stream));
  GG_HIP(hipStreamSynchronize(g->stream));
  if (h & 1u) {
    g->mirror_ok = false;
    g->values_asymmetric = g->value_dtype != GG_VALUES_NONE;
    g->eid.reset();
    return GG_OK;
  }
  if (h & 2u)
    return set_error(GG_ERR_INVALID, "CSR is not symmetric (an edge u->v has no v->u); pass GG_GRAPH_DIRECTED");
  g->mirror_ok = true;
  g->values_asymmetric = (h & 4u) != 0;
  if (!g->values_asymmetric) g->eid.reset();
  return GG_OK;
}

int lbs_buffers(gg_graph *g, GASengine::LbsBuffers *b) {
  const size_t n = (size_t)g->n;
  GG_TRY(g->frontier[0].ensure(4 * n));
  GG_TRY(g->frontier[1].ensure(4 * n));
  GG_TRY(g->scan_local.ensure(4 * n));
  GG_TRY(g->partials.ensure(4 * (n / GASengine::kCountTile + 2)));
  GG_TRY(g->parts.ensure(4 * GASengine::lbs_partition_entries(g->n, g->m)));
  GG_TRY(g->info.ensure(sizeof(GASengine::LbsInfo)));
  GG_TRY(g->counters.ensure(GASengine::kCounterWords * sizeof(unsigned long long)));
  b->scan_local = g->scan_local.as<int32_t>();
  b->partials = g->partials.as<int32_t>();
  b->info = g->info.as<GASengine::LbsInfo>();
  b->parts = g->parts.as<int32_t>();
  b->edge_counter = g->counters.as<unsigned long long>() + 2;
  b->expand_grid = g->opt.expand_grid;
  b->vt = g->opt.lbs_tile / GASengine::kExpandBlock;
  return GG_OK;
}

// ------------------------------------------------------------- profiler --
EventProfiler::~EventProfiler() { reset(); }

hipEvent_t EventProfiler::take() {
  if (!pool_.empty()) {
    hipEvent_t e = pool_.back();
    pool_.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

void EventProfiler::begin(const char *) {
  if (!enabled) return;
  open_ = take();
  (void)hipEventRecord(open_, stream);
}

void EventProfiler::end(const char *name) {
  if (!enabled || !open_) return;
  hipEvent_t b = take();
  (void)hipEventRecord(b, stream);
  pending_.push_back({name, open_, b, 0.0});
  open_ = nullptr;
}

void EventProfiler::add_launch_bytes(const char *name, double bytes) {
  if (!enabled) return;
  table[name].bytes += bytes;
  for (auto it = pending_.rbegin(); it != pending_.rend(); ++it)
    if (it->name == name) {
      it->bytes += bytes;
      return;
    }
}

void EventProfiler::flush() {
  for (auto &p : pending_) {
    float ms = 0.f;
    (void)hipEventSynchronize(p.b);
    (void)hipEventElapsedTime(&ms, p.a, p.b);
    Entry &e = table[p.name];
    e.launches++;
    e.ms += ms;
    if (launches.size() < kMaxLaunches) launches.push_back({p.name, (double)ms, p.bytes});
    pool_.push_back(p.a);
    pool_.push_back(p.b);
  }
  pending_.clear();
}

void EventProfiler::reset() {
  flush();
  table.clear();
  launches.clear();
  for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
  pool_.clear();
}

}  // namespace gg

// ------------------------------------------------------------------ ABI --
extern "C" {

const char *gg_last_error(void) { return g_err; }

int gg_version(void) { return 10000; }

int gg_device_count(int32_t *count) {
  gg::clear_error();
  GG_CHECK_ARG(count, "count is NULL");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) c = 0;
  *count = c;
  return GG_OK;
}

int gg_default_options(gg_options *o) {
  GG_CHECK_ARG(o, "out is NULL");
  o->threshold = -1;  // per-algorithm default: BFS 0, SSSP/CC 10000, PR 1000
  o->wave_degree = 128;
  o->block_degree = 1024;
  o->iter_num = INT_MAX;
  o->bfs_alpha = 14;
  o->bfs_beta = 24;
  o->lbs_tile = GASengine::kLbsTile;
  o->sssp_delta = 0;
  o->pull_alpha = 0;
  o->expand_grid = 2048;
  o->gas_expand_grid = 8192;
  o->xcd_slices = -1;
  return GG_OK;
}

int gg_graph_create(int32_t n, int64_t m, const int32_t *row_offsets, const int32_t *col_indices, const void *values,
                    int32_t value_dtype, uint32_t flags, gg_graph **out) {
  gg::clear_error();
  GG_CHECK_ARG(out, "out is NULL");
  *out = nullptr;
  GG_CHECK_ARG(n >= 1, "n must be >= 1 (got %d)", n);
  GG_CHECK_ARG(m >= 0 && m < INT32_MAX, "m must be in [0, 2^31) (got %lld)", (long long)m);
  GG_CHECK_ARG(row_offsets && (m == 0 || col_indices), "row_offsets/col_indices is NULL");
  GG_CHECK_ARG(value_dtype == GG_VALUES_NONE || value_dtype == GG_VALUES_I32 || value_dtype == GG_VALUES_F32,
               "bad value_dtype %d", value_dtype);
  GG_CHECK_ARG(value_dtype == GG_VALUES_NONE || values, "values is NULL but value_dtype is set");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return gg::set_error(GG_ERR_NO_DEVICE, "no HIP device");

  const bool dev_in = flags & GG_GRAPH_DEVICE_PTRS;
  gg_graph *g = new (std::nothrow) gg_graph();
  if (!g) return gg::set_error(GG_ERR_OOM, "host allocation failed");
  std::unique_ptr<gg_graph, int (*)(gg_graph *)> guard(g, gg_graph_destroy);
  GG_HIP(hipGetDevice(&g->device));
  GG_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
  g->own_stream = true;
  g->prof.stream = g->stream;
  g->n = n;
  g->m = m;
  g->directed = (flags & GG_GRAPH_DIRECTED) ? 1 : 0;
  g->value_dtype = value_dtype;
  gg_default_options(&g->opt);
  GG_HIP(hipHostMalloc((void **)&g->h_pinned, (GASengine::kCounterSlots + GASengine::kCounterWords) * sizeof(int64_t)));
  GG_HIP(hipHostGetDevicePointer((void **)&g->h_dev, g->h_pinned, 0));

  const hipMemcpyKind kind = dev_in ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  GG_TRY(g->ro.ensure(4 * ((size_t)n + 1)));
  GG_TRY(g->col.ensure(4 * (size_t)std::max<int64_t>(m, 1)));
  GG_HIP(hipMemcpy(g->ro.p, row_offsets, 4 * ((size_t)n + 1), kind));
  if (m) GG_HIP(hipMemcpy(g->col.p, col_indices, 4 * (size_t)m, kind));
  if (value_dtype != GG_VALUES_NONE) {
    GG_TRY(g->val.ensure(4 * (size_t)std::max<int64_t>(m, 1)));
    if (m) GG_HIP(hipMemcpy(g->val.p, values, 4 * (size_t)m, kind));
  }
  // Validate the CSR on the device before any kernel indexes with it.
  {
    GG_TRY(g->counters.ensure(GASengine::kCounterWords * sizeof(unsigned long long)));
    int32_t *bad = reinterpret_cast<int32_t *>(g->counters.as<unsigned long long>());
    const int32_t init[4] = {0, INT32_MAX, 0, INT32_MAX};
    GG_HIP(hipMemcpyAsync(bad, init, sizeof(init), hipMemcpyHostToDevice, g->stream));
    hipLaunchKernelGGL(GASengine::k_validate_csr<0>, dim3(GASengine::grid_for(std::max<int64_t>(m, n + 1), 256, 16384)),
                       dim3(256), 0, g->stream, g->ro.as<int32_t>(), g->col.as<int32_t>(), n, m, bad);
    GG_HIP(hipGetLastError());
    int32_t h[4];
    GG_HIP(hipMemcpyAsync(h, bad, sizeof(h), hipMemcpyDeviceToHost, g->stream));
    GG_HIP(hipStreamSynchronize(g->stream));
    GG_CHECK_ARG(h[0] == 0, "row_offsets must start at 0, not decrease and end at m (first violation at %d)", h[1]);
    GG_CHECK_ARG(h[2] == 0, "%d column indices out of range (first at slot %d)", h[2], h[3]);
  }
  if (flags & GG_GRAPH_REORDER) {
    gg::DevBuf ro0, col0, val0;
    ro0.swap(g->ro);
    col0.swap(g->col);
    val0.swap(g->val);
    GG_TRY(gg::build_reorder(g, ro0.as<int32_t>(), col0.as<int32_t>(),
                             value_dtype != GG_VALUES_NONE ? val0.p : nullptr));
  }
  GASengine::DeviceGraph &d = g->dg;
  d.nodes = n;
  d.edges = m;
  d.directed = g->directed;
  d.row_offsets = g->ro.as<int32_t>();
  d.column_indices = g->col.as<int32_t>();
  d.edge_values = g->val.p;
  if (g->directed) {
    GG_TRY(g->co.ensure(4 * ((size_t)n + 1)));
    GG_TRY(g->ri.ensure(4 * (size_t)std::max<int64_t>(m, 1)));
    GG_TRY(g->eid.ensure(4 * (size_t)std::max<int64_t>(m, 1)));
    GG_HIP(GASengine::device_csr_to_csc(n, m, g->ro.as<int32_t>(), g->col.as<int32_t>(), g->co.as<int32_t>(),
                                        g->ri.as<int32_t>(), g->eid.as<int32_t>(), g->stream));
    d.column_offsets = g->co.as<int32_t>();
    d.row_indices = g->ri.as<int32_t>();
    d.csc_edge_ids = g->eid.as<int32_t>();
  } else {
    d.column_offsets = d.row_offsets;
    d.row_indices = d.column_indices;
    d.csc_edge_ids = nullptr;
    if (m) GG_TRY(gg::check_undirected(g));
  }
  *out = guard.release();
  return GG_OK;
}

int gg_graph_destroy(gg_graph *g) {
  if (!g) return GG_OK;
  int cur = -1;
  const int dev = g->device;
  (void)hipGetDevice(&cur);
  if (cur != dev) (void)hipSetDevice(dev);
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  for (auto &p : g->prog_state) p.reset();
  g->prof.reset();
  if (g->h_pinned) (void)hipHostFree(g->h_pinned);
  if (g->own_stream && g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
  if (cur >= 0 && cur != dev) (void)hipSetDevice(cur);
  return GG_OK;
}

int gg_graph_set_stream(gg_graph *g, void *stream) {
  gg::clear_error();
  GG_CHECK_ARG(g, "graph is NULL");
  GG_TRY(gg::activate(g));
  GG_HIP(hipStreamSynchronize(g->stream));
  if (stream) {
    if (g->own_stream) GG_HIP(hipStreamDestroy(g->stream));
    g->stream = (hipStream_t)stream;
    g->own_stream = false;
  } else if (!g->own_stream) {
    GG_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    g->own_stream = true;
  }
  g->prof.stream = g->stream;
  return GG_OK;
}

int gg_graph_set_options(gg_graph *g, const gg_options *o) {
  gg::clear_error();
  GG_CHECK_ARG(g && o, "NULL argument");
  GG_CHECK_ARG(o->wave_degree >= 1 && o->block_degree >= o->wave_degree, "need 1 <= wave_degree <= block_degree");
  GG_CHECK_ARG(o->iter_num >= 1, "iter_num must be >= 1");
  GG_CHECK_ARG(o->bfs_alpha >= 1 && o->bfs_beta >= 1, "bfs_alpha/bfs_beta must be >= 1");
  GG_CHECK_ARG(GASengine::lbs_tile_supported(o->lbs_tile), "lbs_tile must be 512, 1024 or 2048");
  GG_CHECK_ARG(o->expand_grid >= 64 && o->expand_grid <= 65536, "expand_grid must be in [64, 65536]");
  GG_CHECK_ARG(o->gas_expand_grid >= 64 && o->gas_expand_grid <= 65536, "gas_expand_grid must be in [64, 65536]");
  GG_CHECK_ARG(o->xcd_slices >= -1 && o->xcd_slices <= 1, "xcd_slices must be -1, 0 or 1");
  GG_CHECK_ARG(o->sssp_delta >= 0, "sssp_delta must be >= 0");
  GG_CHECK_ARG(o->pull_alpha >= 0, "pull_alpha must be >= 0");
  g->opt = *o;
  return GG_OK;
}

int gg_graph_device_csc(const gg_graph *g, const int32_t **column_offsets, const int32_t **row_indices,
                        const int32_t **csc_edge_ids) {
  gg::clear_error();
  GG_CHECK_ARG(g && column_offsets && row_indices && csc_edge_ids, "NULL argument");
  *column_offsets = g->dg.column_offsets;
  *row_indices = g->dg.row_indices;
  *csc_edge_ids = g->directed ? g->dg.csc_edge_ids : nullptr;
  return GG_OK;
}

int gg_graph_get_options(const gg_graph *g, gg_options *o) {
  GG_CHECK_ARG(g && o, "NULL argument");
  *o = g->opt;
  return GG_OK;
}

int gg_graph_update_values(gg_graph *g, const void *values, uint32_t flags) {
  gg::clear_error();
  GG_CHECK_ARG(g && values, "NULL argument");
  GG_CHECK_ARG(g->value_dtype != GG_VALUES_NONE, "graph was created without values");
  GG_TRY(gg::activate(g));
  if (g->reordered)
    GG_TRY(gg::reorder_in_edges(g, values, flags, g->val.p));
  else if (g->m)
    GG_HIP(hipMemcpyAsync(g->val.p, values, 4 * (size_t)g->m,
                          (flags & GG_OUT_DEVICE) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, g->stream));
  GG_HIP(hipStreamSynchronize(g->stream));
  g->pr_cache_valid = false;
  g->values_version++;
  if (!g->directed && g->m) GG_TRY(gg::check_undirected(g));  // value symmetry may have changed
  return GG_OK;
}

int gg_graph_info(const gg_graph *g, int32_t *n, int64_t *m, int32_t *directed, int32_t *value_dtype) {
  GG_CHECK_ARG(g, "graph is NULL");
  if (n) *n = g->n;
  if (m) *m = g->m;
  if (directed) *directed = g->directed;
  if (value_dtype) *value_dtype = g->value_dtype;
  return GG_OK;
}

int gg_graph_device_csr(const gg_graph *g, const int32_t **row_offsets, const int32_t **col_indices,
                        const void **values) {
  GG_CHECK_ARG(g, "graph is NULL");
  if (row_offsets) *row_offsets = g->dg.row_offsets;
  if (col_indices) *col_indices = g->dg.column_indices;
  if (values) *values = g->dg.edge_values;
  return GG_OK;
}

int gg_graph_vertex_order(const gg_graph *g, const int32_t **new_of_old, const int32_t **old_of_new) {
  gg::clear_error();
  GG_CHECK_ARG(g && new_of_old && old_of_new, "NULL argument");
  *new_of_old = g->reordered ? g->re_new.as<int32_t>() : nullptr;
  *old_of_new = g->reordered ? g->re_order.as<int32_t>() : nullptr;
  return GG_OK;
}

int gg_graph_synchronize(gg_graph *g) {
  gg::clear_error();
  GG_CHECK_ARG(g, "graph is NULL");
  GG_TRY(gg::activate(g));
  GG_HIP(hipStreamSynchronize(g->stream));
  return GG_OK;
}

int gg_profile_reset(gg_graph *g) {
  GG_CHECK_ARG(g, "graph is NULL");
  g->prof.reset();
  return GG_OK;
}

int gg_profile_get(gg_graph *g, gg_kernel_stat *out, int32_t max_entries, int32_t *count) {
  GG_CHECK_ARG(g && count, "NULL argument");
  g->prof.flush();
  int32_t i = 0;
  for (auto &kv : g->prof.table) {
    if (out && i < max_entries) {
      std::memset(&out[i], 0, sizeof(gg_kernel_stat));
      std::strncpy(out[i].name, kv.first.c_str(), sizeof(out[i].name) - 1);
      out[i].launches = kv.second.launches;
      out[i].total_ms = kv.second.ms;
      out[i].alg_bytes = kv.second.bytes;
    }
    i++;
  }
  *count = i;
  return GG_OK;
}

int gg_profile_launches(gg_graph *g, gg_launch_stat *out, int32_t max_entries, int32_t *count) {
  gg::clear_error();
  GG_CHECK_ARG(g && count, "NULL argument");
  g->prof.flush();
  int32_t i = 0;
  for (auto &l : g->prof.launches) {
    if (out && i < max_entries) {
      std::memset(&out[i], 0, sizeof(gg_launch_stat));
      std::strncpy(out[i].name, l.name.c_str(), sizeof(out[i].name) - 1);
      out[i].ms = l.ms;
      out[i].alg_bytes = l.bytes;
    }
    i++;
  }
  *count = i;
  return GG_OK;
}

}  // extern "C"
