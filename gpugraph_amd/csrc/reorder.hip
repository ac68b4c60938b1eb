// Vertex reordering by degree (GG_GRAPH_REORDER).
//
// The random gathers of the traversals and sweeps (dist[v] probes of the
// SSSP push, rank / contribution gathers of PageRank) hit mostly the
// high-degree vertices.  Renumbering vertices by degree, highest first, packs
// those hot entries into the first few megabytes of every vertex array, where
// L2 and the memory-side cache hold them.  The graph is stored renumbered;
// every entry point maps its inputs (sources, seeds, teleport vectors, edge
// values) into the internal numbering and its outputs back, so callers see
// their own ids throughout.  CC labels stay the smallest original id of each
// component (a component's smallest new id is not its smallest old id, so
// the labels are recomputed from the components).
//
// Built once on the device at graph creation:
//   order[i]  old id of new vertex i: vertices by out-degree descending, ties
//             by old id (for directed graphs out-degree ranks the vertices
//             whose values the PageRank pull gathers most; a host relabel by
//             out-degree sped PageRank up 1.21x, by in-degree 1.08x)
//   new_id[v] new id of old vertex v
//   emap[k]   old CSR slot of new slot k; rows keep their edges sorted by new
//             column id
#include <hip/hip_runtime.h>

#include <climits>
#include <cstring>
#include <vector>

#include "internal.h"

namespace gg {
namespace {

using namespace GASengine;

// key = ~out-degree (an ascending sort puts the highest degree first), id = v.
__global__ void k_degree_keys(const int32_t *__restrict__ ro, int32_t n, uint32_t *__restrict__ key,
                              int32_t *__restrict__ id) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t d = (uint32_t)(ro[v + 1] - ro[v]);
    key[v] = ~d;
    id[v] = (int32_t)v;
  }
}

__global__ void k_invert(const int32_t *__restrict__ order, int32_t n, int32_t *__restrict__ new_id) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    new_id[order[i]] = (int32_t)i;
}

// keys[e] = new_row << 32 | new_col, slot[e] = e (one wave per old row).
__global__ void k_edge_keys(const int32_t *__restrict__ ro, const int32_t *__restrict__ col,
                            const int32_t *__restrict__ new_id, int32_t n, uint64_t *__restrict__ keys,
                            int32_t *__restrict__ slot) {
  const int64_t waves = (int64_t)gridDim.x * (blockDim.x / kWave);
  for (int64_t u = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; u < n; u += waves) {
    const uint64_t hi = (uint64_t)(uint32_t)new_id[u] << 32;
    for (int32_t e = ro[u] + lane_id(); e < ro[u + 1]; e += kWave) {
      keys[e] = hi | (uint32_t)new_id[col[e]];
      slot[e] = e;
    }
  }
}

__global__ void k_unpack_cols(const uint64_t *__restrict__ keys, int64_t m, int32_t *__restrict__ col) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += (int64_t)gridDim.x * blockDim.x)
    col[k] = (int32_t)(uint32_t)keys[k];
}

// ro[v] = first sorted slot whose row >= v.
__global__ void k_rows_from_keys(const uint64_t *__restrict__ keys, int64_t m, int32_t n, int32_t *__restrict__ ro) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v <= n; v += (int64_t)gridDim.x * blockDim.x) {
    int64_t lo = 0, hi = m;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if ((int64_t)(keys[mid] >> 32) < v)
        lo = mid + 1;
      else
        hi = mid;
    }
    ro[v] = (int32_t)lo;
  }
}

template <class T>
__global__ void k_gather(const T *__restrict__ in, const int32_t *__restrict__ idx, int64_t count, T *__restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = in[idx[i]];
}

template <class T>
__global__ void k_scatter(const T *__restrict__ in, const int32_t *__restrict__ idx, int64_t count,
                          T *__restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    out[idx[i]] = in[i];
}

// CC: smallest old id of every component (indexed by the component's label
// in new ids), then the old-order output.  Most vertices share one label (the
// giant component), so each wave first folds the lanes that agree with its
// first lane, and every atomic is preceded by a plain read that skips it once
// the stored minimum is already smaller.
__global__ void k_cc_min_old(const int32_t *__restrict__ label, const int32_t *__restrict__ order, int32_t n,
                             int32_t *__restrict__ minold) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
    const int64_t u = base + threadIdx.x;
    const int32_t lab = u < n ? label[u] : -1;
    const int32_t old = u < n ? order[u] : INT_MAX;
    const int32_t lab0 = __shfl(lab, 0);
    const bool same = lab == lab0;
    int32_t m = same ? old : INT_MAX;
    for (int off = 32; off > 0; off >>= 1) m = min(m, __shfl_xor(m, off));
    if (lane_id() == 0 && lab0 >= 0 && m < minold[lab0]) atomicMin(&minold[lab0], m);
    if (!same && lab >= 0 && old < minold[lab]) atomicMin(&minold[lab], old);
  }
}

__global__ void k_cc_old_labels(const int32_t *__restrict__ label, const int32_t *__restrict__ minold,
                                const int32_t *__restrict__ new_id, int32_t n, int32_t *__restrict__ out) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x)
    out[v] = minold[label[new_id[v]]];
}

}  // namespace

int build_reorder(gg_graph *g, const int32_t *ro0, const int32_t *col0, const void *val0) {
  const int32_t n = g->n;
  const int64_t m = g->m;
  hipStream_t s = g->stream;
  GG_TRY(g->ro.ensure(4 * ((size_t)n + 1)));
  GG_TRY(g->col.ensure(4 * (size_t)std::max<int64_t>(m, 1)));
  GG_TRY(g->re_order.ensure(4 * (size_t)n));
  GG_TRY(g->re_new.ensure(4 * (size_t)n));
  GG_TRY(g->re_emap.ensure(4 * (size_t)std::max<int64_t>(m, 1)));
  // 1. order: vertices by out-degree descending (stable).
  {
    DevBuf key_in, key_out, id_in;
    GG_TRY(key_in.ensure(4 * (size_t)n));
    GG_TRY(key_out.ensure(4 * (size_t)n));
    GG_TRY(id_in.ensure(4 * (size_t)n));
    hipLaunchKernelGGL(k_degree_keys, dim3(grid_for(n, 256, 16384)), dim3(256), 0, s, ro0, n, key_in.as<uint32_t>(),
                       id_in.as<int32_t>());
    GG_HIP(hipGetLastError());
    GG_HIP((GASengine::radix_sort_pairs<uint32_t, int32_t>(key_in.as<uint32_t>(), key_out.as<uint32_t>(),
                                                            id_in.as<int32_t>(), g->re_order.as<int32_t>(), n, 0, 32,
                                                            s)));
    hipLaunchKernelGGL(k_invert, dim3(grid_for(n, 256, 16384)), dim3(256), 0, s, g->re_order.as<int32_t>(), n,
                       g->re_new.as<int32_t>());
    GG_HIP(hipGetLastError());
    GG_HIP(hipStreamSynchronize(s));
  }
  // 2. renumbered CSR: sort the slots by (new row, new column).
  if (m > 0) {
    DevBuf keys_in, keys_out, slot_in;
    GG_TRY(keys_in.ensure(8 * (size_t)m));
    GG_TRY(keys_out.ensure(8 * (size_t)m));
    GG_TRY(slot_in.ensure(4 * (size_t)m));
    hipLaunchKernelGGL(k_edge_keys, dim3(grid_for((int64_t)n * kWave, 256, 16384)), dim3(256), 0, s, ro0, col0,
                       g->re_new.as<int32_t>(), n, keys_in.as<uint64_t>(), slot_in.as<int32_t>());
    GG_HIP(hipGetLastError());
    int bits = 1;
    while (bits < 31 && ((int64_t)1 << bits) < n) bits++;
    GG_HIP((GASengine::radix_sort_pairs<uint64_t, int32_t>(keys_in.as<uint64_t>(), keys_out.as<uint64_t>(),
                                                            slot_in.as<int32_t>(), g->re_emap.as<int32_t>(), m, 0,
                                                            32 + bits, s)));
    hipLaunchKernelGGL(k_unpack_cols, dim3(grid_for(m, 256, 16384)), dim3(256), 0, s, keys_out.as<uint64_t>(), m,
                       g->col.as<int32_t>());
    hipLaunchKernelGGL(k_rows_from_keys, dim3(grid_for((int64_t)n + 1, 256, 16384)), dim3(256), 0, s,
                       keys_out.as<uint64_t>(), m, n, g->ro.as<int32_t>());
    GG_HIP(hipGetLastError());
    if (val0) {
      GG_TRY(g->val.ensure(4 * (size_t)m));
      hipLaunchKernelGGL(k_gather<uint32_t>, dim3(grid_for(m, 256, 16384)), dim3(256), 0, s,
                         static_cast<const uint32_t *>(val0), g->re_emap.as<int32_t>(), m, g->val.as<uint32_t>());
      GG_HIP(hipGetLastError());
    }
    GG_HIP(hipStreamSynchronize(s));
  } else {
    GG_HIP(fill_async<int32_t>(g->ro.as<int32_t>(), 0, (int64_t)n + 1, s));
    if (val0) GG_TRY(g->val.ensure(4));
  }
  g->h_new_id.resize((size_t)n);
  GG_HIP(hipMemcpyAsync(g->h_new_id.data(), g->re_new.p, 4 * (size_t)n, hipMemcpyDeviceToHost, s));
  GG_HIP(hipStreamSynchronize(s));
  g->reordered = true;
  return GG_OK;
}

int32_t map_vertex(const gg_graph *g, int32_t v) { return g->reordered ? g->h_new_id[(size_t)v] : v; }

namespace {
// The mapping of reorder_out_vertices, enqueued only.
int enqueue_out_vertices(gg_graph *g, const void *in, void *dst, uint32_t flags) {
  const int32_t n = g->n;
  hipStream_t s = g->stream;
  uint32_t *out = static_cast<uint32_t *>(dst);
  if (!(flags & GG_OUT_DEVICE)) {
    GG_TRY(g->re_stage.ensure(4 * (size_t)n));
    out = g->re_stage.as<uint32_t>();
  }
  hipLaunchKernelGGL(k_gather<uint32_t>, dim3(grid_for(n, 256, 16384)), dim3(256), 0, s,
                     static_cast<const uint32_t *>(in), g->re_new.as<int32_t>(), (int64_t)n, out);
  GG_HIP(hipGetLastError());
  if (!(flags & GG_OUT_DEVICE)) GG_HIP(hipMemcpyAsync(dst, out, 4 * (size_t)n, hipMemcpyDeviceToHost, s));
  return GG_OK;
}
}  // namespace

// 4-byte vertex results in internal order (device) -> caller order (host or
// device per flags).
int reorder_out_vertices(gg_graph *g, const void *in, void *dst, uint32_t flags) {
  if (!dst) return GG_OK;
  GG_TRY(enqueue_out_vertices(g, in, dst, flags));
  GG_HIP(hipStreamSynchronize(g->stream));
  return GG_OK;
}

int enqueue_handoff(gg_graph *g, const void *result) {
  if (!g->handoff.armed || !g->handoff.dst) return GG_OK;
  return enqueue_out_vertices(g, result, g->handoff.dst, g->handoff.flags);
}

// 4-byte vertex inputs in caller order (host or device per flags) -> internal
// order in buf.
int reorder_in_vertices(gg_graph *g, const void *src, uint32_t flags, DevBuf &buf, const void **out) {
  const int32_t n = g->n;
  hipStream_t s = g->stream;
  const void *in = src;
  DevBuf host_copy;
  if (!(flags & GG_OUT_DEVICE)) {
    GG_TRY(host_copy.ensure(4 * (size_t)n));
    GG_HIP(hipMemcpyAsync(host_copy.p, src, 4 * (size_t)n, hipMemcpyHostToDevice, s));
    in = host_copy.p;
  }
  GG_TRY(buf.ensure(4 * (size_t)n));
  hipLaunchKernelGGL(k_gather<uint32_t>, dim3(grid_for(n, 256, 16384)), dim3(256), 0, s,
                     static_cast<const uint32_t *>(in), g->re_order.as<int32_t>(), (int64_t)n, buf.as<uint32_t>());
  GG_HIP(hipGetLastError());
  GG_HIP(hipStreamSynchronize(s));
  *out = buf.p;
  return GG_OK;
}

// 4-byte edge values in caller (original CSR) order -> internal order.
int reorder_in_edges(gg_graph *g, const void *src, uint32_t flags, void *dst_dev) {
  const int64_t m = g->m;
  if (m == 0) return GG_OK;
  hipStream_t s = g->stream;
  const void *in = src;
  DevBuf host_copy;
  if (!(flags & GG_OUT_DEVICE)) {
    GG_TRY(host_copy.ensure(4 * (size_t)m));
    GG_HIP(hipMemcpyAsync(host_copy.p, src, 4 * (size_t)m, hipMemcpyHostToDevice, s));
    in = host_copy.p;
  }
  hipLaunchKernelGGL(k_gather<uint32_t>, dim3(grid_for(m, 256, 16384)), dim3(256), 0, s,
                     static_cast<const uint32_t *>(in), g->re_emap.as<int32_t>(), m, static_cast<uint32_t *>(dst_dev));
  GG_HIP(hipGetLastError());
  GG_HIP(hipStreamSynchronize(s));
  return GG_OK;
}

// 4-byte edge results in internal order (device) -> caller order.
int reorder_out_edges(gg_graph *g, const void *in, void *dst, uint32_t flags) {
  const int64_t m = g->m;
  if (!dst || m == 0) return GG_OK;
  hipStream_t s = g->stream;
  uint32_t *out = static_cast<uint32_t *>(dst);
  DevBuf stage;
  if (!(flags & GG_OUT_DEVICE)) {
    GG_TRY(stage.ensure(4 * (size_t)m));
    out = stage.as<uint32_t>();
  }
  hipLaunchKernelGGL(k_scatter<uint32_t>, dim3(grid_for(m, 256, 16384)), dim3(256), 0, s,
                     static_cast<const uint32_t *>(in), g->re_emap.as<int32_t>(), m, out);
  GG_HIP(hipGetLastError());
  if (!(flags & GG_OUT_DEVICE)) GG_HIP(hipMemcpyAsync(dst, out, 4 * (size_t)m, hipMemcpyDeviceToHost, s));
  GG_HIP(hipStreamSynchronize(s));
  return GG_OK;
}

// CC labels in internal ids (device) -> smallest original id per component,
// caller order.
int reorder_out_cc(gg_graph *g, const int32_t *label, int32_t *dst, uint32_t flags) {
  if (!dst) return GG_OK;
  const int32_t n = g->n;
  hipStream_t s = g->stream;
  DevBuf minold;
  GG_TRY(minold.ensure(4 * (size_t)n));
  GG_HIP(fill_async<int32_t>(minold.as<int32_t>(), INT_MAX, n, s));
  hipLaunchKernelGGL(k_cc_min_old, dim3(grid_for(n, 256, 16384)), dim3(256), 0, s, label, g->re_order.as<int32_t>(),
                     n, minold.as<int32_t>());
  int32_t *out = dst;
  if (!(flags & GG_OUT_DEVICE)) {
    GG_TRY(g->re_stage.ensure(4 * (size_t)n));
    out = g->re_stage.as<int32_t>();
  }
  hipLaunchKernelGGL(k_cc_old_labels, dim3(grid_for(n, 256, 16384)), dim3(256), 0, s, label, minold.as<int32_t>(),
                     g->re_new.as<int32_t>(), n, out);
  GG_HIP(hipGetLastError());
  if (!(flags & GG_OUT_DEVICE)) GG_HIP(hipMemcpyAsync(dst, out, 4 * (size_t)n, hipMemcpyDeviceToHost, s));
  GG_HIP(hipStreamSynchronize(s));
  return GG_OK;
}

}  // namespace gg
