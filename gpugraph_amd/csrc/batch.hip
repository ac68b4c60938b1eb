// Batched instances: up to 128 SSSP sources and 64 personalised-PageRank
// seeds per pass, solved together (gg_batch_run).
//
// Replaces the reference drivers' sequential per-source loop (sssp.cu:389-404,
// bfs.cu:865-881: one EnactIterativeSearch per source).
//
// SSSP: dist is vertex-major, dist[v][128] (512 B per vertex); lane l of a
//   wave owns instances l and l + 64, so one edge (u, v, w) is relaxed for
//   all instances with two coalesced 256 B row accesses instead of 128
//   random 4-byte probes.  Frontier = union of the instances' frontiers: a
//   vertex queue deduplicated by a bitmap, plus a 128-bit active mask per
//   vertex.  Per iteration:
//     k_bs_count   degrees of the queue (workgroup scan; clears the queue's
//                  bits in the union bitmap)
//     k_lbs_scan   partials -> offsets, total edges
//     k_bs_partition  merge-path tile boundaries (items + edges)
//     k_bs_relax   per tile, each wave relaxes a contiguous quarter of its
//                  edges for all instances, four edges' row loads in flight:
//                  atomicMin on dist[v][i] after a plain pre-check, improved
//                  instances OR-ed into the next active mask (at most four
//                  atomics per wave per edge), new vertices staged in LDS and
//                  appended with one atomic per tile
//     k_bs_clear   clears the active masks just consumed
//     k_bs_swap    one thread: swaps queue / bitmap / mask buffers, counts,
//                  sets done when the next queue is empty
// PPR: R is vertex-major, R[v][64] (lane = instance).  One sweep is a pull
//   over CSC in-edges: rows up to kLongRow in one wave (Neumaier-compensated
//   FP32 per lane, the source row R[u][:] read as one 256 B access), longer
//   rows in chunks whose partial sums a second kernel folds in order;
//   teleport one-hot on seed[l]; FP64 dangling mass and L1 residual per
//   instance; k_ppr_finalize stops when every instance is below tol.
// Every kernel reads its buffers and sizes from a device-side state block
// and returns at once when its part is done, so the whole iteration is a
// fixed launch sequence: it is captured once into a hipGraph (kIters
// iterations per replay) and replayed until both parts are done.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstring>

#include "gasengine/rowreduce.h"
#include "internal.h"

namespace gg {

int ensure_transition(gg_graph *g);

namespace {

using namespace GASengine;
constexpr int32_t kInf = 100000000;
constexpr int kSsspMax = 128;
constexpr int kPprMax = 64;
constexpr int kLongRow = 512;     // PPR rows longer than this are chunked
constexpr int kIters = 8;         // iterations per graph replay

struct BatchState {
  // SSSP part
  int32_t *q_cur, *q_nxt;          // union frontier queues
  uint32_t *ub_cur, *ub_nxt;       // union bitmaps (dedup of the queues)
  uint32_t *act_cur, *act_nxt;     // [n][4] active-instance masks
  int32_t cur_size, nxt_size;
  int32_t sssp_iters, sssp_done;
  unsigned long long edges;        // edge relaxations x instances-groups (edges expanded)
  // PPR part
  float *r_cur, *r_nxt;            // [n][64]
  int32_t ppr_iters, ppr_done;
  double resid_last_max;
  double dang_cur[kPprMax];
  // Per-instance sums of a sweep, striped over kStripes copies (stride
  // kPprMax doubles; see GASengine::block_counter_add), folded by
  // k_ppr_finalize.
  double dang_acc[kStripes * kPprMax];
  double resid_acc[kStripes * kPprMax];
};

// ---------------------------------------------------------------- SSSP --

__global__ void k_bs_init(const int32_t *__restrict__ src, int32_t k, int32_t *dist, BatchState *st) {
  const int i = threadIdx.x;
  if (i >= k) return;
  const int32_t v = src[i];
  dist[(int64_t)v * kSsspMax + i] = 0;
  atomicOr(&st->act_cur[(int64_t)v * 4 + (i >> 5)], 1u << (i & 31));
  if (bit_test_and_set(st->ub_cur, v)) {
    const int32_t at = atomicAdd(&st->cur_size, 1);
    st->q_cur[at] = v;
  }
}

__global__ void __launch_bounds__(kCountBlock)
k_bs_count(BatchState *st, const int32_t *__restrict__ ro, int32_t *__restrict__ scan_local,
           int32_t *__restrict__ partials) {
  __shared__ int32_t smem[kCountBlock / kWave + 1];
  if (st->sssp_done) return;
  const int32_t nf = st->cur_size;
  const int32_t *q = st->q_cur;
  uint32_t *ub = st->ub_cur;
  for (int64_t b = blockIdx.x; b * kCountTile < nf; b += gridDim.x) {
    const int64_t first = b * kCountTile + (int64_t)threadIdx.x * kCountIpt;
    int32_t deg[kCountIpt];
    int32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kCountIpt; k++) {
      deg[k] = 0;
      if (first + k < nf) {
        const int32_t v = q[first + k];
        ub[v >> 5] = 0u;  // every set bit there belongs to a queued vertex
        deg[k] = ro[v + 1] - ro[v];
      }
      sum += deg[k];
    }
    int32_t total;
    int32_t ex = block_exclusive_scan<kCountBlock>(sum, smem, &total);
#pragma unroll
    for (int k = 0; k < kCountIpt; k++) {
      if (first + k < nf) scan_local[first + k] = ex;
      ex += deg[k];
    }
    if (threadIdx.x == 0) partials[b] = total;
  }
}

// Wrapper so the scan reads the size from the state block.
__global__ void __launch_bounds__(kScanBlock)
k_bs_scan(BatchState *st, int32_t *__restrict__ partials, LbsInfo *__restrict__ info) {
  __shared__ int32_t smem[kScanBlock / kWave + 1];
  if (st->sssp_done) return;
  const int32_t nf = st->cur_size;
  const int32_t nb = (int32_t)(((int64_t)nf + kCountTile - 1) / kCountTile);
  int32_t carry = 0;
  for (int32_t base = 0; base < nb; base += kScanBlock) {
    int32_t i = base + threadIdx.x;
    int32_t v = i < nb ? partials[i] : 0;
    int32_t tot;
    int32_t ex = block_exclusive_scan<kScanBlock>(v, smem, &tot);
    if (i < nb) partials[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) {
    info->edges = carry;
    info->path = (int64_t)nf + carry;
  }
}

// Merge-path tile boundaries of the current queue (k_lbs_partition with the
// queue size taken from the state block).
__global__ void __launch_bounds__(256)
k_bs_partition(BatchState *st, const int32_t *__restrict__ scan_local, const int32_t *__restrict__ partials,
               const LbsInfo *__restrict__ info, int32_t *__restrict__ parts) {
  if (st->sssp_done) return;
  const int32_t nf = st->cur_size;
  const int32_t ne = info->edges;
  const int64_t path = info->path;
  const int64_t ntiles = (path + kLbsTile - 1) / kLbsTile;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t <= ntiles; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t d = t * kLbsTile < path ? t * kLbsTile : path;
    parts[t] = lbs_path_search(scan_local, partials, d, nf, ne);
  }
}

// Relax one edge (u -> v, w) for every instance: lane l owns instances l and
// l + 64.  m0/m1: u's active words for this lane; du0/du1: u's distances;
// dv0/dv1: v's distances as pre-read.  Returns true when v must join the
// next frontier (some instance improved); ORs the improved instances into
// v's next mask (two words per improved half, one lane each).
__device__ __forceinline__ bool relax_row(int32_t v, int32_t w, uint32_t m0, uint32_t m1, int32_t du0, int32_t du1,
                                          int32_t dv0, int32_t dv1, int32_t *dist, uint32_t *act_nxt) {
  const int lane = lane_id();
  const int64_t vrow = (int64_t)v * kSsspMax;
  bool imp0 = false, imp1 = false;
  if ((m0 >> (lane & 31)) & 1u) {
    const int64_t nd = (int64_t)du0 + w;
    if (nd < kInf && nd < dv0) imp0 = (int32_t)nd < atomicMin(&dist[vrow + lane], (int32_t)nd);
  }
  if ((m1 >> (lane & 31)) & 1u) {
    const int64_t nd = (int64_t)du1 + w;
    if (nd < kInf && nd < dv1) imp1 = (int32_t)nd < atomicMin(&dist[vrow + 64 + lane], (int32_t)nd);
  }
  const uint64_t b0 = ballot(imp0), b1 = ballot(imp1);
  if (!(b0 | b1)) return false;
  const uint32_t word = lane == 0 ? (uint32_t)b0 : lane == 1 ? (uint32_t)(b0 >> 32)
                      : lane == 2 ? (uint32_t)b1 : lane == 3 ? (uint32_t)(b1 >> 32) : 0u;
  if (lane < 4 && word) atomicOr(&act_nxt[(int64_t)v * 4 + lane], word);
  return true;
}

// Tile-based relax: the merge path of (queue items, edges) in tiles of
// kLbsTile (same partition as k_lbs_expand); each wave takes a contiguous
// quarter of the tile's edges and relaxes them for all instances, four edges'
// loads in flight at a time.  New frontier vertices are staged in LDS and
// appended with one atomic per tile.
__global__ void __launch_bounds__(kExpandBlock)
k_bs_relax(BatchState *st, const int32_t *__restrict__ ro, const int32_t *__restrict__ col,
           const int32_t *__restrict__ wt, int32_t *__restrict__ dist, const int32_t *__restrict__ scan_local,
           const int32_t *__restrict__ partials, const LbsInfo *__restrict__ info, const int32_t *__restrict__ parts) {
  constexpr int TILE = kLbsTile;
  constexpr int VT = kExpandVt;
  constexpr int W = kExpandBlock / kWave;
  __shared__ int32_t s_offs[TILE + 1];
  __shared__ int32_t s_vert[TILE + 1];
  __shared__ int32_t s_begin[TILE + 1];
  __shared__ int32_t s_slot[TILE];
  __shared__ int32_t s_out[TILE];
  __shared__ int32_t s_nout, s_obase;
  if (st->sssp_done) return;
  const int32_t *queue = st->q_cur;
  const uint32_t *act = st->act_cur;
  uint32_t *act_nxt = st->act_nxt;
  uint32_t *ub_nxt = st->ub_nxt;
  const int lane = lane_id(), wid = threadIdx.x / kWave;
  const int64_t path = info->path;
  for (int64_t t = blockIdx.x; t * TILE < path; t += gridDim.x) {
    const int64_t d0 = t * TILE;
    const int64_t d1 = d0 + TILE < path ? d0 + TILE : path;
    const int32_t i0 = parts[t], i1 = parts[t + 1];
    const int32_t j0 = (int32_t)(d0 - i0), j1 = (int32_t)(d1 - i1);
    const int32_t ni = i1 - i0, ne = j1 - j0;
    for (int k = threadIdx.x; k <= ni; k += kExpandBlock) {
      const int32_t i = i0 - 1 + k;
      if (i >= 0) {
        const int32_t v = queue[i];
        s_vert[k] = v;
        s_offs[k] = lbs_offset(scan_local, partials, i);
        s_begin[k] = ro[v];
      } else {
        s_vert[k] = 0;
        s_offs[k] = 0;
        s_begin[k] = 0;
      }
    }
    if (threadIdx.x == 0) s_nout = 0;
    __syncthreads();
    {
      const int32_t dl = (int32_t)threadIdx.x * VT;
      if (dl < ni + ne) {
        int32_t lo = dl - ne > 0 ? dl - ne : 0, hi = dl < ni ? dl : ni;
        while (lo < hi) {
          int32_t mid = (lo + hi) >> 1;
          if (s_offs[mid + 1] <= j0 + (dl - mid) - 1)
            lo = mid + 1;
          else
            hi = mid;
        }
        int32_t li = lo, lj = dl - lo;
#pragma unroll
        for (int k = 0; k < VT; k++) {
          if (li + lj >= ni + ne) break;
          if (li < ni && s_offs[li + 1] <= j0 + lj) {
            li++;
          } else {
            s_slot[lj] = li;
            lj++;
          }
        }
      }
    }
    __syncthreads();
    const int32_t per = (ne + W - 1) / W;
    const int32_t kb = wid * per, ke = min(ne, kb + per);
    for (int32_t k0 = kb; k0 < ke; k0 += 4) {
      int32_t vv[4], ww[4], uu[4], du0[4], du1[4], dv0[4], dv1[4];
      uint32_t m0[4], m1[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int32_t k = min(k0 + q, ke - 1);  // tail: repeat the last edge, masked below
        const int32_t s = s_slot[k];
        const int32_t u = s_vert[s];
        const int32_t slot = s_begin[s] + (j0 + k - s_offs[s]);
        vv[q] = col[slot];
        ww[q] = wt[slot];
        uu[q] = u;
        m0[q] = act[(int64_t)u * 4 + (lane >> 5)];
        m1[q] = act[(int64_t)u * 4 + 2 + (lane >> 5)];
      }
      // Row loads only for active instances: a sparse mask touches a few
      // cache lines instead of both 256 B halves of two rows.  The target
      // row is read with plain (cacheable) loads: a stale, larger distance
      // only lets an unneeded atomicMin through, which returns the truth.
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int64_t urow = (int64_t)uu[q] * kSsspMax, vrow = (int64_t)vv[q] * kSsspMax;
        du0[q] = dv0[q] = du1[q] = dv1[q] = kInf;
        if ((m0[q] >> (lane & 31)) & 1u) {
          du0[q] = dist[urow + lane];
          dv0[q] = dist[vrow + lane];
        }
        if ((m1[q] >> (lane & 31)) & 1u) {
          du1[q] = dist[urow + 64 + lane];
          dv1[q] = dist[vrow + 64 + lane];
        }
      }
#pragma unroll
      for (int q = 0; q < 4; q++) {
        if (k0 + q >= ke) break;
        if (relax_row(vv[q], ww[q], m0[q], m1[q], du0[q], du1[q], dv0[q], dv1[q], dist, act_nxt) && lane == 0 &&
            bit_test_and_set(ub_nxt, vv[q]))
          s_out[atomicAdd(&s_nout, 1)] = vv[q];
      }
    }
    __syncthreads();
    const int32_t nout = s_nout;
    if (nout) {
      if (threadIdx.x == 0) s_obase = atomicAdd(&st->nxt_size, nout);
      __syncthreads();
      for (int32_t k = threadIdx.x; k < nout; k += kExpandBlock) st->q_nxt[s_obase + k] = s_out[k];
    }
    __syncthreads();
  }
}

__global__ void k_bs_clear(BatchState *st) {
  if (st->sssp_done) return;
  const int32_t nf = st->cur_size;
  uint32_t *act = st->act_cur;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nf; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t u = st->q_cur[i];
    act[u * 4 + 0] = 0u;
    act[u * 4 + 1] = 0u;
    act[u * 4 + 2] = 0u;
    act[u * 4 + 3] = 0u;
  }
}

__global__ void k_bs_swap(BatchState *st, const LbsInfo *info, int32_t max_iter) {
  if (st->sssp_done) return;
  st->edges += (unsigned long long)info->edges;
  int32_t *q = st->q_cur;
  st->q_cur = st->q_nxt;
  st->q_nxt = q;
  uint32_t *b = st->ub_cur;
  st->ub_cur = st->ub_nxt;
  st->ub_nxt = b;
  uint32_t *a = st->act_cur;
  st->act_cur = st->act_nxt;
  st->act_nxt = a;
  st->cur_size = st->nxt_size;
  st->nxt_size = 0;
  st->sssp_iters++;
  if (st->cur_size == 0 || st->sssp_iters >= max_iter) st->sssp_done = 1;
}

// Vertex-major rows -> instance-major output: in[row(v)][KW] -> out[i][v] for
// i < k, row(v) = perm[v] on a renumbered graph (v in caller ids), else v.  A
// workgroup moves 64 vertices through LDS (row-contiguous reads, 256 B per
// instance writes; the +1 pad keeps both LDS passes conflict-free).
template <typename T, int KW>
__global__ void __launch_bounds__(256)
k_rows_to_instances(const T *__restrict__ in, int32_t n, int32_t k, const int32_t *__restrict__ perm,
                    T *__restrict__ out) {
  __shared__ T tile[64][KW + 1];
  for (int64_t v0 = (int64_t)blockIdx.x * 64; v0 < n; v0 += (int64_t)gridDim.x * 64) {
    const int nv = n - v0 < 64 ? (int)(n - v0) : 64;
    for (int idx = threadIdx.x; idx < nv * KW; idx += 256) {
      const int r = idx / KW, j = idx % KW;
      const int64_t row = perm ? perm[v0 + r] : v0 + r;
      tile[r][j] = in[row * KW + j];
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < k * 64; idx += 256) {
      const int i = idx / 64, v = idx % 64;
      if (v < nv) out[(int64_t)i * n + v0 + v] = tile[v][i];
    }
    __syncthreads();
  }
}

// ----------------------------------------------------------------- PPR --

struct PprArgs {
  const int32_t *co;
  const int32_t *ri;
  const float *pcsc;     // weighted P per CSC slot, or nullptr (1/outdeg)
  const float *inv_out;
  const int32_t *seeds;  // [64]
  int32_t nseeds;
  int32_t n;
  float d;
  double inv_n;
};

template <int BLOCK>
__device__ __forceinline__ void lane_add_f64(double *acc_per_instance, double v, double *s_red) {
  // s_red: [BLOCK/64][64]; sums the same lane of every wave, one atomic per instance.
  const int w = threadIdx.x / kWave, l = lane_id();
  s_red[w * kWave + l] = v;
  __syncthreads();
  if (threadIdx.x < kWave) {
    double t = 0.0;
    for (int j = 0; j < BLOCK / kWave; j++) t += s_red[j * kWave + threadIdx.x];
    if (t != 0.0) atomicAdd(stripe_of<kPprMax>(acc_per_instance) + threadIdx.x, t);
  }
  __syncthreads();
}

// acc += R[u][l] * P(u -> v) over in-edge slots [b, e) in order, four
// source rows (256 B each, one per edge for the whole wave) in flight.
__device__ __forceinline__ void ppr_row_sum(const PprArgs &a, const float *rc, int32_t b, int32_t e, int l, KSum &acc) {
  int32_t k = b;
  for (; k + 3 < e; k += 4) {
    float x[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int32_t u = a.ri[k + q];
      const float p = a.pcsc ? a.pcsc[k + q] : a.inv_out[u];
      x[q] = rc[(int64_t)u * kPprMax + l] * p;
    }
#pragma unroll
    for (int q = 0; q < 4; q++) acc.add(x[q]);
  }
  for (; k < e; k++) {
    const int32_t u = a.ri[k];
    const float p = a.pcsc ? a.pcsc[k] : a.inv_out[u];
    acc.add(rc[(int64_t)u * kPprMax + l] * p);
  }
}

__device__ __forceinline__ void ppr_finish_vertex(const PprArgs &a, BatchState *st, int32_t v, float sum,
                                                  const float *rc, float *rn, double &resid, double &dang) {
  const int l = lane_id();
  const float dn = (float)(st->dang_cur[l] * a.inv_n);
  const float tele = (l < a.nseeds && a.seeds[l] == v) ? (1.0f - a.d) : 0.0f;
  const float r = a.d * (sum + dn) + tele;
  rn[(int64_t)v * kPprMax + l] = r;
  resid += fabs((double)r - (double)rc[(int64_t)v * kPprMax + l]);
  if (a.inv_out[v] == 0.0f) dang += r;
}

// Rows up to kLongRow: one wave per vertex, lanes are instances.
__global__ void __launch_bounds__(kRowBlock)
k_ppr_sweep(PprArgs a, BatchState *st) {
  __shared__ double s_red[kRowBlock];
  if (st->ppr_done) return;
  const float *rc = st->r_cur;
  float *rn = st->r_nxt;
  const int l = lane_id();
  double resid = 0.0, dang = 0.0;
  const int64_t waves = (int64_t)gridDim.x * (kRowBlock / kWave);
  for (int64_t v = (int64_t)blockIdx.x * (kRowBlock / kWave) + threadIdx.x / kWave; v < a.n; v += waves) {
    const int32_t b = a.co[v], e = a.co[v + 1];
    if (e - b > kLongRow) continue;
    KSum acc;
    ppr_row_sum(a, rc, b, e, l, acc);
    ppr_finish_vertex(a, st, (int32_t)v, acc.value(), rc, rn, resid, dang);
  }
  lane_add_f64<kRowBlock>(st->resid_acc, resid, s_red);
  lane_add_f64<kRowBlock>(st->dang_acc, dang, s_red);
}

// Long rows: task t = (long vertex j, chunk c) writes its partial (s, c) per
// lane; chunk_base[j] = first task of vertex j (prefix over the long list).
__global__ void __launch_bounds__(kRowBlock)
k_ppr_long_partial(PprArgs a, BatchState *st, const int32_t *__restrict__ longv, const int32_t *__restrict__ chunk_base,
                   int32_t nlong, int32_t ntasks, float2 *__restrict__ part) {
  if (st->ppr_done) return;
  const float *rc = st->r_cur;
  const int l = lane_id();
  const int64_t waves = (int64_t)gridDim.x * (kRowBlock / kWave);
  for (int64_t t = (int64_t)blockIdx.x * (kRowBlock / kWave) + threadIdx.x / kWave; t < ntasks; t += waves) {
    int32_t lo = 0, hi = nlong;  // vertex j with chunk_base[j] <= t < chunk_base[j+1]
    while (hi - lo > 1) {
      const int32_t mid = (lo + hi) >> 1;
      if (chunk_base[mid] <= t)
        lo = mid;
      else
        hi = mid;
    }
    const int32_t v = longv[lo];
    const int32_t c = (int32_t)t - chunk_base[lo];
    const int32_t b = a.co[v] + c * kLongRow;
    const int32_t e = min(a.co[v + 1], b + kLongRow);
    KSum acc;
    ppr_row_sum(a, rc, b, e, l, acc);
    part[t * kPprMax + l] = make_float2(acc.s, acc.c);
  }
}

__global__ void __launch_bounds__(kRowBlock)
k_ppr_long_finish(PprArgs a, BatchState *st, const int32_t *__restrict__ longv, const int32_t *__restrict__ chunk_base,
                  int32_t nlong, const float2 *__restrict__ part) {
  __shared__ double s_red[kRowBlock];
  if (st->ppr_done) return;
  const float *rc = st->r_cur;
  float *rn = st->r_nxt;
  const int l = lane_id();
  double resid = 0.0, dang = 0.0;
  const int64_t waves = (int64_t)gridDim.x * (kRowBlock / kWave);
  for (int64_t j = (int64_t)blockIdx.x * (kRowBlock / kWave) + threadIdx.x / kWave; j < nlong; j += waves) {
    KSum acc;
    for (int32_t t = chunk_base[j]; t < chunk_base[j + 1]; t++) {
      const float2 p = part[(int64_t)t * kPprMax + l];
      acc.merge(p.x, p.y);
    }
    ppr_finish_vertex(a, st, longv[j], acc.value(), rc, rn, resid, dang);
  }
  lane_add_f64<kRowBlock>(st->resid_acc, resid, s_red);
  lane_add_f64<kRowBlock>(st->dang_acc, dang, s_red);
}

// 64 threads: close the iteration for every instance.
__global__ void k_ppr_finalize(BatchState *st, int32_t nseeds, double tol, int32_t max_iter) {
  __shared__ double s_max[kPprMax];
  if (st->ppr_done) return;
  const int l = threadIdx.x;
  double r = 0.0, dg = 0.0;
  for (int q = 0; q < kStripes; q++) {
    r += st->resid_acc[q * kPprMax + l];
    dg += st->dang_acc[q * kPprMax + l];
    st->resid_acc[q * kPprMax + l] = 0.0;
    st->dang_acc[q * kPprMax + l] = 0.0;
  }
  if (l >= nseeds) r = 0.0;
  st->dang_cur[l] = dg;
  s_max[l] = r;
  __syncthreads();
  if (l == 0) {
    double mx = 0.0;
    for (int j = 0; j < kPprMax; j++) mx = fmax(mx, s_max[j]);
    st->resid_last_max = mx;
    float *t = st->r_cur;
    st->r_cur = st->r_nxt;
    st->r_nxt = t;
    st->ppr_iters++;
    if (mx <= tol || st->ppr_iters >= max_iter) st->ppr_done = 1;
  }
}

// R0[v][l] = one-hot on seed[l]; dangling mass per instance.
__global__ void k_ppr_init(PprArgs a, BatchState *st, float *r0) {
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < (int64_t)a.n * kPprMax;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = idx / kPprMax;
    const int l = (int)(idx % kPprMax);
    const float r = (l < a.nseeds && a.seeds[l] == v) ? 1.0f : 0.0f;
    r0[idx] = r;
    if (r != 0.0f && a.inv_out[v] == 0.0f) atomicAdd(&st->dang_cur[l], (double)r);
  }
}

__global__ void k_bs_state_init(BatchState *st, int32_t *q0, int32_t *q1, uint32_t *ub0, uint32_t *ub1,
                                uint32_t *a0, uint32_t *a1, float *r0, float *r1, int32_t sssp_on, int32_t ppr_on) {
  st->q_cur = q0;
  st->q_nxt = q1;
  st->ub_cur = ub0;
  st->ub_nxt = ub1;
  st->act_cur = a0;
  st->act_nxt = a1;
  st->cur_size = 0;
  st->nxt_size = 0;
  st->sssp_iters = 0;
  st->sssp_done = sssp_on ? 0 : 1;
  st->edges = 0;
  st->r_cur = r0;
  st->r_nxt = r1;
  st->ppr_iters = 0;
  st->ppr_done = ppr_on ? 0 : 1;
  st->resid_last_max = INFINITY;
  for (int i = 0; i < kPprMax; i++) st->dang_cur[i] = 0.0;
  for (int i = 0; i < kStripes * kPprMax; i++) st->dang_acc[i] = st->resid_acc[i] = 0.0;
}

}  // namespace

// One pass: <= 128 SSSP sources, <= 64 PPR seeds.
static int batch_pass(gg_graph *g, const int32_t *h_src, int32_t ns, int32_t *dist_out, const int32_t *h_seeds,
                      int32_t np, float *rank_out, float damping, double tol, int32_t max_iter, uint32_t flags,
                      gg_stats *acc) {
  const int32_t n = g->n;
  const int64_t m = g->m;
  hipStream_t s = g->stream;
  const size_t words = ((size_t)n + 31) / 32;
  // Buffers: batch[0] dist, [1] act x2, [2] union bitmaps x2, [3] queues x2,
  //          [4] R x2, [5] state, [6] sources/seeds, [7] long-row partials.
  if (ns) {
    GG_TRY(g->batch[0].ensure(4 * (size_t)n * kSsspMax));
    GG_TRY(g->batch[1].ensure(2 * 16 * (size_t)n));
    GG_TRY(g->batch[2].ensure(2 * 4 * words));
    GG_TRY(g->batch[3].ensure(2 * 4 * (size_t)n));
  }
  if (np) {
    GG_TRY(g->batch[4].ensure(2 * 4 * (size_t)n * kPprMax));
    GG_TRY(ensure_transition(g));
  }
  GG_TRY(g->batch[5].ensure(sizeof(BatchState)));
  GG_TRY(g->batch[6].ensure(4 * (kSsspMax + kPprMax)));
  LbsBuffers lb;
  GG_TRY(lbs_buffers(g, &lb));
  BatchState *st = g->batch[5].as<BatchState>();
  int32_t *d_src = g->batch[6].as<int32_t>(), *d_seed = d_src + kSsspMax;
  int32_t *dist = g->batch[0].as<int32_t>();
  uint32_t *act = g->batch[1].as<uint32_t>(), *ub = g->batch[2].as<uint32_t>();
  int32_t *qs = g->batch[3].as<int32_t>();
  float *R = g->batch[4].as<float>();
  if (ns) GG_HIP(hipMemcpyAsync(d_src, h_src, 4 * (size_t)ns, hipMemcpyHostToDevice, s));
  if (np) GG_HIP(hipMemcpyAsync(d_seed, h_seeds, 4 * (size_t)np, hipMemcpyHostToDevice, s));

  // Long PPR rows (> kLongRow in-edges) from the CSC mid/large bins, and
  // their chunk prefix: built on the host once per graph.
  PprArgs pa{};
  int32_t nlong = 0, ntasks = 0;
  int32_t *d_long = nullptr, *d_cbase = nullptr;
  float2 *d_part = nullptr;
  if (np) {
    std::vector<int32_t> co((size_t)n + 1);
    GG_HIP(hipMemcpyAsync(co.data(), g->dg.column_offsets, 4 * ((size_t)n + 1), hipMemcpyDeviceToHost, s));
    GG_HIP(hipStreamSynchronize(s));
    std::vector<int32_t> lv, cb(1, 0);
    for (int32_t v = 0; v < n; v++) {
      const int32_t dgr = co[v + 1] - co[v];
      if (dgr > kLongRow) {
        lv.push_back(v);
        cb.push_back(cb.back() + (dgr + kLongRow - 1) / kLongRow);
      }
    }
    nlong = (int32_t)lv.size();
    ntasks = cb.back();
    GG_TRY(g->batch[7].ensure(4 * (lv.size() + cb.size()) + 8 * (size_t)std::max(ntasks, 1) * kPprMax + 16));
    d_long = g->batch[7].as<int32_t>();
    d_cbase = d_long + lv.size();
    d_part = reinterpret_cast<float2 *>(
        (reinterpret_cast<uintptr_t>(d_cbase + cb.size()) + 15) & ~uintptr_t(15));
    if (nlong) {
      GG_HIP(hipMemcpyAsync(d_long, lv.data(), 4 * lv.size(), hipMemcpyHostToDevice, s));
      GG_HIP(hipMemcpyAsync(d_cbase, cb.data(), 4 * cb.size(), hipMemcpyHostToDevice, s));
    }
    pa.co = g->dg.column_offsets;
    pa.ri = g->dg.row_indices;
    pa.pcsc = g->value_dtype == GG_VALUES_F32 ? g->pr[3].as<float>() : nullptr;
    pa.inv_out = g->pr[0].as<float>();
    pa.seeds = d_seed;
    pa.nseeds = np;
    pa.n = n;
    pa.d = damping;
    pa.inv_n = 1.0 / (double)n;
  }

  ScopedEvents ev;
  GG_HIP(ev.create());
  hipEvent_t e0 = ev.e0, e1 = ev.e1;
  GG_HIP(hipEventRecord(e0, s));
  {
    TimedRegion t(g->prof, "batch.init");
    hipLaunchKernelGGL(k_bs_state_init, dim3(1), dim3(1), 0, s, st, qs, qs + n, ub, ub + words, act, act + 4 * (size_t)n,
                       R, R + (size_t)n * kPprMax, ns > 0, np > 0);
    if (ns) {
      GG_HIP(fill_async<int32_t>(dist, kInf, (int64_t)n * kSsspMax, s));
      GG_HIP(fill_async<uint32_t>(act, 0u, 8 * (int64_t)n, s));
      GG_HIP(fill_async<uint32_t>(ub, 0u, 2 * (int64_t)words, s));
      hipLaunchKernelGGL(k_bs_init, dim3(1), dim3(kSsspMax), 0, s, d_src, ns, dist, st);
    }
    if (np)
      hipLaunchKernelGGL(k_ppr_init, dim3(grid_for((int64_t)n * kPprMax, 256, 16384)), dim3(256), 0, s, pa, st, R);
    GG_HIP(hipGetLastError());
  }
  // Capture kIters iterations of both parts into one graph.
  ScopedGraph cg;
  GG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  for (int it = 0; it < kIters; it++) {
    if (ns) {
      hipLaunchKernelGGL(k_bs_count, dim3(grid_for(n, kCountTile, kCountGridMax)), dim3(kCountBlock), 0, s, st,
                         g->dg.row_offsets, lb.scan_local, lb.partials);
      hipLaunchKernelGGL(k_bs_scan, dim3(1), dim3(kScanBlock), 0, s, st, lb.partials, lb.info);
      hipLaunchKernelGGL(k_bs_partition, dim3(grid_for(((int64_t)n + m) / kLbsTile + 2, 256, 4096)), dim3(256), 0, s,
                         st, lb.scan_local, lb.partials, lb.info, lb.parts);
      hipLaunchKernelGGL(k_bs_relax, dim3(lb.expand_grid), dim3(kExpandBlock), 0, s, st, g->dg.row_offsets,
                         g->dg.column_indices, static_cast<const int32_t *>(g->dg.edge_values), dist, lb.scan_local,
                         lb.partials, lb.info, lb.parts);
      hipLaunchKernelGGL(k_bs_clear, dim3(grid_for(n, 256, 2048)), dim3(256), 0, s, st);
      hipLaunchKernelGGL(k_bs_swap, dim3(1), dim3(1), 0, s, st, lb.info, INT_MAX);  // SSSP runs to quiescence
    }
    if (np) {
      hipLaunchKernelGGL(k_ppr_sweep, dim3(grid_for((int64_t)n * kWave, kRowBlock, 8192)), dim3(kRowBlock), 0, s, pa,
                         st);
      if (nlong) {
        hipLaunchKernelGGL(k_ppr_long_partial, dim3(grid_for((int64_t)ntasks * kWave, kRowBlock, 4096)),
                           dim3(kRowBlock), 0, s, pa, st, d_long, d_cbase, nlong, ntasks, d_part);
        hipLaunchKernelGGL(k_ppr_long_finish, dim3(grid_for((int64_t)nlong * kWave, kRowBlock, 4096)), dim3(kRowBlock),
                           0, s, pa, st, d_long, d_cbase, nlong, d_part);
      }
      hipLaunchKernelGGL(k_ppr_finalize, dim3(1), dim3(kPprMax), 0, s, st, np, tol, max_iter);
    }
  }
  hipError_t cap = hipStreamEndCapture(s, &cg.graph);
  if (cap != hipSuccess) return set_error(GG_ERR_HIP, "graph capture failed: %s", hipGetErrorString(cap));
  GG_HIP(hipGraphInstantiate(&cg.exec, cg.graph, nullptr, nullptr, 0));
  struct H {
    int32_t sssp_iters, sssp_done;
    unsigned long long edges;
  };
  int replays = 0;
  for (;;) {
    {
      TimedRegion t(g->prof, "batch.replay");
      GG_HIP(hipGraphLaunch(cg.exec, s));
    }
    replays++;
    GG_HIP(hipMemcpyAsync(g->h_pinned, &st->sssp_iters, sizeof(H), hipMemcpyDeviceToHost, s));
    GG_HIP(hipMemcpyAsync(g->h_pinned + 8, &st->ppr_iters, 8, hipMemcpyDeviceToHost, s));
    GG_HIP(hipMemcpyAsync(g->h_pinned + 10, &st->resid_last_max, 8, hipMemcpyDeviceToHost, s));
    GG_HIP(hipStreamSynchronize(s));
    const int sd = g->h_pinned[1], pd = g->h_pinned[9];
    if (sd && pd) break;
  }
  H h;
  std::memcpy(&h, g->h_pinned, sizeof(H));
  const int32_t ppr_iters = g->h_pinned[8];
  double resid;
  std::memcpy(&resid, g->h_pinned + 10, 8);
  const int32_t *perm = g->reordered ? g->re_new.as<int32_t>() : nullptr;
  {
    TimedRegion t(g->prof, "batch.output");
    if (ns && dist_out) {
      int32_t *o = dist_out;
      if (!(flags & GG_OUT_DEVICE)) {
        GG_TRY(g->vtx[2].ensure(4 * (size_t)n * ns));
        o = g->vtx[2].as<int32_t>();
      }
      hipLaunchKernelGGL((k_rows_to_instances<int32_t, kSsspMax>), dim3(grid_for(n, 64, 16384)), dim3(256), 0, s,
                         dist, n, ns, perm, o);
      GG_HIP(hipGetLastError());
      if (!(flags & GG_OUT_DEVICE))
        GG_HIP(hipMemcpyAsync(dist_out, o, 4 * (size_t)n * ns, hipMemcpyDeviceToHost, s));
    }
    if (np && rank_out) {
      const float *Rf = (ppr_iters & 1) ? R + (size_t)n * kPprMax : R;
      float *o = rank_out;
      if (!(flags & GG_OUT_DEVICE)) {
        GG_TRY(g->vtx[1].ensure(4 * (size_t)n * np));
        o = g->vtx[1].as<float>();
      }
      hipLaunchKernelGGL((k_rows_to_instances<float, kPprMax>), dim3(grid_for(n, 64, 16384)), dim3(256), 0, s, Rf,
                         n, np, perm, o);
      GG_HIP(hipGetLastError());
      if (!(flags & GG_OUT_DEVICE))
        GG_HIP(hipMemcpyAsync(rank_out, o, 4 * (size_t)n * np, hipMemcpyDeviceToHost, s));
    }
  }
  GG_HIP(hipEventRecord(e1, s));
  GG_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  acc->iterations = std::max(acc->iterations, h.sssp_iters);
  acc->pull_iterations = std::max(acc->pull_iterations, ppr_iters);
  acc->edges_inspected += (int64_t)h.edges + (int64_t)ppr_iters * m;
  acc->gpu_ms += ms;
  acc->residual = std::max(acc->residual, np ? resid : 0.0);
  // Bytes: SSSP 512 B per expanded edge (dist row of the target) + 8 B (col, w);
  // PPR per sweep 264 B per in-edge (source row + id + weight) + 512 B per vertex.
  acc->alg_bytes += (double)h.edges * 520.0 + (double)ppr_iters * (264.0 * m + 512.0 * n);
  if (np && resid > tol)
    return set_error(GG_ERR_NOT_CONVERGED, "batched PPR: residual %.3e > tol %.3e after %d sweeps", resid, tol,
                     ppr_iters);
  (void)replays;
  return GG_OK;
}

}  // namespace gg

extern "C" int gg_batch_run(gg_graph *g, const gg_batch_spec *spec, uint32_t flags, gg_stats *stats) {
  using namespace gg;
  gg::clear_error();
  GG_CHECK_ARG(g && spec, "graph/spec is NULL");
  GG_CHECK_ARG(spec->num_sssp >= 0 && spec->num_ppr >= 0 && spec->num_bfs >= 0, "negative instance count");
  GG_CHECK_ARG(spec->num_bfs == 0 || (spec->bfs_sources && spec->bfs_depth), "BFS instances need sources and output");
  GG_CHECK_ARG(spec->num_sssp == 0 || (spec->sssp_sources && g->value_dtype == GG_VALUES_I32),
               "SSSP instances need sources and int32 weights");
  GG_CHECK_ARG(spec->num_ppr == 0 || spec->ppr_seeds, "PPR instances need seeds");
  GG_CHECK_ARG(spec->num_ppr == 0 || (spec->damping >= 0.f && spec->damping < 1.f && spec->max_iter >= 1),
               "bad damping / max_iter");
  for (int32_t i = 0; i < spec->num_sssp; i++)
    GG_CHECK_ARG(spec->sssp_sources[i] >= 0 && spec->sssp_sources[i] < g->n, "source %d out of range", i);
  for (int32_t i = 0; i < spec->num_ppr; i++)
    GG_CHECK_ARG(spec->ppr_seeds[i] >= 0 && spec->ppr_seeds[i] < g->n, "seed %d out of range", i);
  for (int32_t i = 0; i < spec->num_bfs; i++)
    GG_CHECK_ARG(spec->bfs_sources[i] >= 0 && spec->bfs_sources[i] < g->n, "BFS source %d out of range", i);
  GG_TRY(activate(g));
  if (spec->num_sssp) GG_TRY(require_nonnegative_weights(g, "batched SSSP"));
  g->prof.enabled = (flags & GG_INSTRUMENT) != 0;
  gg_stats acc{};
  const int32_t max_iter = spec->max_iter > 0 ? spec->max_iter : INT_MAX;
  int32_t done_s = 0, done_p = 0;
  const size_t n = (size_t)g->n;
  // Sources and seeds in the graph's internal numbering.
  std::vector<int32_t> src_int, seed_int;
  if (spec->num_sssp) src_int.assign(spec->sssp_sources, spec->sssp_sources + spec->num_sssp);
  if (spec->num_ppr) seed_int.assign(spec->ppr_seeds, spec->ppr_seeds + spec->num_ppr);
  for (auto &v : src_int) v = map_vertex(g, v);
  for (auto &v : seed_int) v = map_vertex(g, v);
  while (done_s < spec->num_sssp || done_p < spec->num_ppr) {
    const int32_t ns = std::min(kSsspMax, spec->num_sssp - done_s);
    const int32_t np = std::min(kPprMax, spec->num_ppr - done_p);
    int32_t *dout = spec->sssp_dist ? spec->sssp_dist + done_s * n : nullptr;
    float *rout = spec->ppr_rank ? spec->ppr_rank + done_p * n : nullptr;
    GG_TRY(batch_pass(g, src_int.empty() ? nullptr : src_int.data() + done_s, ns, dout,
                      seed_int.empty() ? nullptr : seed_int.data() + done_p, np,
                      rout, spec->damping, spec->tol_l1, max_iter, flags, &acc));
    done_s += ns;
    done_p += np;
  }
  if (spec->num_bfs) GG_TRY(batch_bfs(g, spec->bfs_sources, spec->num_bfs, spec->bfs_depth, flags, &acc));
  finish_profile(g);
  if (stats) *stats = acc;
  return GG_OK;
}
