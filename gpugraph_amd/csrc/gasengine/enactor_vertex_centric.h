// EnactorVertexCentric: the GAS iteration driver (reference:
// GASengine/enactor_vertex_centric.cuh:96-5072; its single-GPU sequence is
// gather -> apply -> post_apply -> expand -> contract, :3454-3771).
//
// One iteration on the problem's stream:
//   k_gas_ctl       stop test and LBS-or-tiered choice on the device (GasCtl)
//   [gather]        frontier gather over in- and/or out-edges, load-balanced
//                   by merge path (gather.h: LBS count / scan / partition,
//                   k_lbs_gather, k_lbs_gather_fixup)
//   expand, LBS     k_lbs_count (apply + post_apply fused as the per-item
//                   prologue), k_lbs_scan, k_lbs_partition, k_lbs_expand
//                   (expand_edge + contract + dedup + enqueue per edge)
//   expand, tiered  k_frontier_prologue (apply + post_apply) when the program
//                   has them, k_tiered_expand
//   k_gas_order     a next frontier of at least n/64 vertices rewritten in
//                   vertex order from its dedup bitmap
// Every kernel reads its frontier size from GasCtl, which holds 0 for the
// path not taken and for all kernels once the run is over.
//
// Device-driven (default): kLoopIters iterations are captured once into a
// hipGraph (both expansion paths present, one gated off per iteration),
// cached on the problem with the arguments it was captured with, and
// replayed until the loop state says done: one host read per replay instead
// of the reference's host read and synchronisation per iteration
// (enactor_vertex_centric.cuh:3801-3804).  Host-driven (a KernelTimer is
// set, or DEBUG output is on): the same kernels one iteration at a time, only
// the chosen expansion launched, with per-iteration records.
//
// post_apply runs right after apply in the same thread: the barrier the
// reference places between them (kernel boundary) only matters for programs
// whose post_apply reads another vertex's apply output, and none of the
// vertex programs of this path do.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstring>
#include <optional>
#include <vector>

#include "csr_problem.h"
#include "gather.h"

namespace GASengine {

// Optional per-kernel timing hook (the reference's INSTRUMENT flag collects
// per-CTA clocks, b40c/util/kernel_runtime_stats.cuh:40-130).
struct KernelTimer {
  virtual void begin(const char *name) = 0;
  virtual void end(const char *name) = 0;
  virtual ~KernelTimer() = default;
};

struct EnactorConfig {
  int32_t wave_degree = 64;
  int32_t block_degree = 4096;
  int32_t expand_grid = 8192;  // workgroups of k_lbs_expand / k_lbs_gather (strided over the tiles)
  int32_t expand_vt = kExpandVt;  // LBS tile = kExpandBlock * expand_vt positions (2, 4 or 8)
  // 0: LBS or tiered by frontier size against the threshold, as the
  // reference chooses.  > 0: a frontier within the threshold that holds a
  // row longer than this is expanded by LBS too (the tiered kernel gives
  // such a row to one workgroup).
  int32_t tier_max_degree = 0;
  bool device_loop = true;     // capture iterations into a hipGraph (off: one host decision per iteration)
};

// Per-iteration record: frontier size in, edges expanded, frontier size out.
struct IterationRecord {
  int32_t frontier;
  int64_t edges;
  int32_t next;
  int lbs;
};

// Iterations per captured replay (even: the ping-pong buffers keep their
// parity across replays).
constexpr int kLoopIters = 8;

template <int DUMMY = 0>
__global__ void k_gas_ctl_init(GasCtl *c, int32_t *sizes, int32_t nf0, int32_t threshold, int32_t iter_num,
                               int32_t tier_max_degree) {
  c->iter = 0;
  c->done = 0;
  c->lbs_size = c->tier_size = c->gather_size = 0;
  c->threshold = threshold;
  c->iter_num = iter_num;
  c->iterations = 0;
  c->lbs_iters = 0;
  c->order_epoch = 0;
  c->tier_max_degree = tier_max_degree;
  c->sum_frontier = 0;
  c->sum_next = 0;
  sizes[0] = nf0;
  sizes[1] = 0;
}

// Start of an iteration whose input frontier is sizes[sel] (listed in
// frontier).  With tier_max_degree set, a frontier of at most
// kCtlBlock * kCtlItems vertices within the threshold has its rows' lengths
// read (offs_a, and offs_b when the program expands both ways; all loads of
// a thread at once) and goes to LBS if one is longer.  Thread 0 reads the
// state and writes it; the other threads only read the frontier.
constexpr int kCtlBlock = 1024;
constexpr int kCtlItems = 16;
template <int DUMMY = 0>
__global__ void __launch_bounds__(kCtlBlock)
k_gas_ctl(GasCtl *c, int32_t *sizes, int sel, const int32_t *__restrict__ frontier,
          const int32_t *__restrict__ offs_a, const int32_t *__restrict__ offs_b) {
  __shared__ int32_t s_nf, s_limit;
  if (threadIdx.x == 0) {
    const int32_t nf = sizes[sel];
    s_nf = c->done ? -1 : nf;
    const bool check = !c->done && nf > 0 && nf <= c->threshold && nf <= kCtlBlock * kCtlItems &&
                       c->iterations < c->iter_num && c->tier_max_degree > 0;
    s_limit = check ? c->tier_max_degree : 0;
  }
  __syncthreads();
  const int32_t nf = s_nf, limit = s_limit;
  if (nf < 0) return;
  int long_row = 0;
  if (limit > 0) {
    int32_t v[kCtlItems];
#pragma unroll
    for (int k = 0; k < kCtlItems; k++) {
      const int32_t i = (int32_t)threadIdx.x + k * kCtlBlock;
      v[k] = i < nf ? frontier[i] : -1;
    }
#pragma unroll
    for (int k = 0; k < kCtlItems; k++) {
      if (v[k] < 0) continue;
      int32_t d = offs_a[v[k] + 1] - offs_a[v[k]];
      if (offs_b) d = max(d, offs_b[v[k] + 1] - offs_b[v[k]]);
      long_row |= d > limit;
    }
  }
  long_row = __syncthreads_or(long_row);
  if (threadIdx.x != 0) return;
  if (c->iterations > 0) c->sum_next += nf;  // the previous iteration's output
  if (nf == 0 || c->iterations >= c->iter_num) {
    c->done = 1;
    c->lbs_size = c->tier_size = c->gather_size = 0;
    return;
  }
  sizes[sel ^ 1] = 0;
  c->iter = c->iterations;
  c->iterations++;
  c->sum_frontier += nf;
  const bool lbs = nf > c->threshold || long_row;
  c->lbs_iters += lbs ? 1 : 0;
  c->lbs_size = lbs ? nf : 0;
  c->tier_size = lbs ? 0 : nf;
  c->gather_size = nf;
}

// The next frontier listed in vertex order from the dedup bitmap that built
// it: always after an LBS expansion (whose edges only mark the bitmap, so no
// workgroup serialises on the queue's size counter), and after a tiered one
// (which appends, in whatever order its waves finish) when the frontier has
// at least min_size vertices; the next iteration then reads its vertex
// state and rows in runs of consecutive vertices.  Each thread counts a run
// of per_thread consecutive words; queue slots come from a single-pass scan
// across the workgroups (block_exclusive_across), and the last workgroup
// writes the size.  Same gating argument as every collect of this engine:
// the state is read at each workgroup's start and written only by the last
// workgroup, after every other one has published its count.  A workgroup's
// ids are staged in LDS (kOrderStage per pass) and stored from there, so
// consecutive lanes store consecutive queue slots.
constexpr int kOrderStage = 8192;
template <int DUMMY = 0>
__global__ void __launch_bounds__(kOrderBlock)
k_gas_order(GasCtl *__restrict__ c, const uint32_t *__restrict__ bits, int32_t words, int32_t per_thread,
            int32_t *__restrict__ next_size, int32_t min_size, int32_t *__restrict__ queue,
            uint64_t *__restrict__ status) {
  if (c->done || (c->lbs_size == 0 && *next_size < min_size)) return;
  const int par = c->order_epoch & 1;
  const int G = gridDim.x, b = blockIdx.x;
  if (threadIdx.x == 0) status[(par ^ 1) * G + b] = 0ull;  // this workgroup's word of the next run
  __shared__ int32_t s_scan[kOrderBlock / kWave + 1];
  const int64_t w0 = ((int64_t)b * kOrderBlock + threadIdx.x) * per_thread;
  int32_t cnt = 0;
  for (int k = 0; k < per_thread; k++)
    if (w0 + k < words) cnt += __popc(bits[w0 + k]);
  int32_t total;
  int32_t pos = block_exclusive_scan<kOrderBlock>(cnt, s_scan, &total);
  const bool last = b == G - 1;
  const uint64_t excl = block_exclusive_across<kOrderBlock>(status + par * G, b, (uint64_t)total, total > 0 || last);
  __shared__ int32_t s_ids[kOrderStage];
  for (int32_t base = 0; base < total; base += kOrderStage) {  // total is the workgroup's own count
    const int32_t lim = min(total - base, kOrderStage);
    if (cnt > 0 && pos < base + lim && pos + cnt > base) {
      int32_t p = pos;
      for (int k = 0; k < per_thread && p < base + lim; k++) {
        if (w0 + k >= words) break;
        uint32_t word = bits[w0 + k];
        while (word && p < base + lim) {
          const int bb = __builtin_ctz(word);
          word &= word - 1;
          if (p >= base) s_ids[p - base] = (int32_t)((w0 + k) * 32 + bb);
          p++;
        }
      }
    }
    __syncthreads();
    for (int32_t k = threadIdx.x; k < lim; k += kOrderBlock) queue[(int64_t)excl + base + k] = s_ids[k];
    __syncthreads();
  }
  if (last && threadIdx.x == 0) {
    *next_size = (int32_t)excl + total;
    c->order_epoch++;
  }
}

// MARK_PREDECESSORS: -1 everywhere, then the initial frontier (the sources)
// as its own predecessors.
template <int DUMMY = 0>
__global__ void k_preds_sources(int32_t *preds, const int32_t *frontier, const int32_t *fsize) {
  const int32_t nf = *fsize;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nf; i += (int64_t)gridDim.x * blockDim.x)
    preds[frontier[i]] = frontier[i];
}

template <class CsrProblemT, class Program, bool INSTRUMENT = false>
class EnactorVertexCentric {
 public:
  EnactorConfig config;
  KernelTimer *timer = nullptr;
  std::vector<IterationRecord> history;  // host-driven runs only
  int64_t total_edges = 0;
  // Totals of the last run (both loop kinds).
  int32_t iterations_run = 0;
  int32_t lbs_iterations = 0;
  int64_t frontier_total = 0, next_total = 0;
  int32_t replays = 0;  // device-driven runs: graph launches
  int verbose = 0;
  std::optional<Config> cfg;  // held when constructed from one (see CsrProblem::cfg)

  explicit EnactorVertexCentric(bool verbose_ = false) : verbose(verbose_) {}
  // Reference constructor (enactor_vertex_centric.cuh:273): the driver's
  // parameters and its DEBUG flag (per-iteration output).
  EnactorVertexCentric(const Config &cfg_, bool DEBUG) : verbose(DEBUG ? 1 : 0), cfg(cfg_) {}
  explicit EnactorVertexCentric(const Config &cfg_) : cfg(cfg_) {}

  // Reference 7-argument form (as called by Algorithms/SSSP/sssp.cu:397-398).
  hipError_t EnactIterativeSearch(CsrProblemT &problem, const int32_t * /*h_row_offsets*/, int directed,
                                  int num_srcs, int *srcs, int iter_num = INT_MAX, int threshold = 0) {
    hipStream_t s = problem.stream;
    Workspace &ws = problem.ws;
    DeviceGraph &g = problem.graph;
    const int32_t n = g.nodes;
    history.clear();
    total_edges = 0;
    iterations_run = lbs_iterations = replays = 0;
    frontier_total = next_total = 0;

    GAS_RETURN_IF(problem.Reset());
    GAS_RETURN_IF(ws.AllocateValues());
    int *keys[3] = {ws.frontier[0], ws.frontier[1], ws.frontier[1]};
    typename Program::MiscType *vals[3];
    for (int i = 0; i < 3; i++) vals[i] = (typename Program::MiscType *)ws.frontier_values[i];
    engine_stream() = s;
    Program::Initialize(directed, n, (int)g.edges, num_srcs, srcs, (int *)g.row_offsets, (int *)g.column_indices,
                        (int *)g.column_offsets, (int *)g.row_indices, (typename Program::DataType *)g.edge_values,
                        problem.vertex_list, problem.edge_list, keys, vals);
    GAS_RETURN_IF(hipGetLastError());
    const int32_t nf0 = Program::srcVertex() == SINGLE ? num_srcs : n;
    hipLaunchKernelGGL(k_gas_ctl_init<>, dim3(1), dim3(1), 0, s, ws.ctl, ws.sizes, nf0, threshold,
                       iter_num < 0 ? 0 : iter_num, config.tier_max_degree);
    GAS_RETURN_IF(fill_async<uint64_t>(ws.order_words, 0ull, 2 * kOrderGridMax, s));
    if (CsrProblemT::kMarkPredecessors) {
      GAS_RETURN_IF(ws.AllocatePreds());
      GAS_RETURN_IF(fill_async<int32_t>(ws.preds, -1, n, s));
      hipLaunchKernelGGL(k_preds_sources<>, dim3(grid_for(nf0, 256, 4096)), dim3(256), 0, s, ws.preds, ws.frontier[0],
                         ws.sizes);
    }
    GAS_RETURN_IF(hipGetLastError());
    if (Program::gatherOverEdges() != NO_GATHER_EDGES)
      GAS_RETURN_IF(ws.EnsureCarry(sizeof(GatherCarry<typename Program::GatherType>) *
                                   (size_t)gather_carry_entries((int64_t)n + g.edges)));

    GasState<Program> st = state(problem);
    if (config.device_loop && !timer && !verbose) return device_loop(problem, st, threshold);
    return host_loop(problem, st, threshold, iter_num);
  }

 private:
  GasState<Program> state(CsrProblemT &problem) {
    Workspace &ws = problem.ws;
    GasState<Program> st;
    std::memset(static_cast<void *>(&st), 0, sizeof(st));  // padding too: the state's bytes key the captured loop
    st.g = problem.graph;
    st.vl = problem.vertex_list;
    st.el = problem.edge_list;
    st.gather_tmp = (typename Program::GatherType *)ws.gather_tmp;
    st.changed = ws.changed;
    st.do_apply = Program::applyOverEdges() != NO_APPLY_VERTICES;
    st.do_post = Program::postApplyOverEdges() != NO_POST_APPLY_VERTICES;
    st.iter_ptr = &ws.ctl->iter;
    st.preds = CsrProblemT::kMarkPredecessors ? ws.preds : nullptr;
    return st;
  }

  // Tiered grid for device-driven loops: enough workgroups for a frontier of
  // `threshold` items (the largest it handles), the kernel strides beyond.
  int tier_grid(int threshold, int32_t n) const {
    const int64_t cap = std::min<int64_t>((int64_t)threshold < 0 ? 0 : threshold, n);
    return grid_for(std::max<int64_t>(cap, 1), kExpandBlock, 2048);
  }

  // The kernels of iteration parity `sel`.  mode: 0 both expansions (gated on
  // the device), 1 LBS only, 2 tiered only.  nf_host sizes grids (n when
  // unknown).
  void launch_gas_ctl(CsrProblemT &problem, int sel, hipStream_t s) {
    Workspace &ws = problem.ws;
    const DeviceGraph &g = problem.graph;
    const ExpandEdges ex = Program::expandOverEdges();
    const int32_t *offs_a = nullptr, *offs_b = nullptr;
    if (ex == EXPAND_OUT_EDGES || ex == EXPAND_ALL_EDGES) offs_a = g.row_offsets;
    if (ex == EXPAND_IN_EDGES || ex == EXPAND_ALL_EDGES) (offs_a ? offs_b : offs_a) = g.column_offsets;
    if (!offs_a) offs_a = g.row_offsets;  // no expansion: the check is off anyway (no long-row case)
    hipLaunchKernelGGL(k_gas_ctl<>, dim3(1), dim3(kCtlBlock), 0, s, ws.ctl, ws.sizes, sel, ws.frontier[sel], offs_a,
                       offs_b);
  }

  void issue_iteration(CsrProblemT &problem, GasState<Program> st, int sel, int mode, int64_t nf_host, int tgrid,
                       hipStream_t s, bool launch_ctl = true) {
    Workspace &ws = problem.ws;
    const DeviceGraph &g = problem.graph;
    const int32_t n = g.nodes;
    GasCtl *c = ws.ctl;
    int32_t *f_in = ws.frontier[sel], *f_out = ws.frontier[sel ^ 1];
    int32_t *sz_out = ws.sizes + (sel ^ 1);
    st.dedup_next = ws.dedup[sel];
    st.dedup_prev = ws.dedup[sel ^ 1];
    if (launch_ctl) launch_gas_ctl(problem, sel, s);
    const int64_t max_path = (int64_t)n + g.edges;
    const ExpandEdges ex = Program::expandOverEdges();
    const GatherEdges ga = Program::gatherOverEdges();

    if (ga != NO_GATHER_EDGES) {
      tb("gather");
      using GC = GatherCarry<typename Program::GatherType>;
      GC *carry = (GC *)ws.gcarry;
      for (int pass = 0; pass < 2; pass++) {
        const bool in = pass == 0;
        if (in && !(ga == GATHER_IN_EDGES || ga == GATHER_ALL_EDGES)) continue;
        if (!in && !(ga == GATHER_OUT_EDGES || ga == GATHER_ALL_EDGES)) continue;
        const bool accum = !in && ga == GATHER_ALL_EDGES;
        GatherRows<Program> gop{st, in ? 1 : 0};
        hipLaunchKernelGGL((k_lbs_count<GatherRows<Program>, false>), dim3(grid_for(nf_host, kCountTile, kCountGridMax)),
                           dim3(kCountBlock), 0, s, gop, f_in, &c->gather_size, ws.scan_local, ws.partials);
        hipLaunchKernelGGL(k_lbs_scan<kScanBlock>, dim3(1), dim3(kScanBlock), 0, s, &c->gather_size, ws.partials,
                           ws.info, (unsigned long long *)nullptr);
        hipLaunchKernelGGL(k_lbs_partition<kGatherTile>, dim3(grid_for(max_path / kGatherTile + 2, 256, 4096)),
                           dim3(256), 0, s, &c->gather_size, ws.scan_local, ws.partials, ws.info, ws.parts);
        if (accum) {
          hipLaunchKernelGGL((k_lbs_gather<Program, true>), dim3(config.expand_grid), dim3(kExpandBlock), 0, s, gop,
                             f_in, &c->gather_size, ws.scan_local, ws.partials, ws.info, ws.parts, carry);
          hipLaunchKernelGGL((k_lbs_gather_fixup<Program, true>), dim3(grid_for(max_path / kGatherTile + 2, 256, 1024)),
                             dim3(256), 0, s, gop, f_in, ws.info, carry);
        } else {
          hipLaunchKernelGGL((k_lbs_gather<Program, false>), dim3(config.expand_grid), dim3(kExpandBlock), 0, s, gop,
                             f_in, &c->gather_size, ws.scan_local, ws.partials, ws.info, ws.parts, carry);
          hipLaunchKernelGGL((k_lbs_gather_fixup<Program, false>), dim3(grid_for(max_path / kGatherTile + 2, 256, 1024)),
                             dim3(256), 0, s, gop, f_in, ws.info, carry);
        }
      }
      te("gather");
    }

    int passes[2];
    int np = 0;
    if (ex == EXPAND_OUT_EDGES || ex == EXPAND_ALL_EDGES) passes[np++] = 0;
    if (ex == EXPAND_IN_EDGES || ex == EXPAND_ALL_EDGES) passes[np++] = 1;
    if (np == 0) {
      // No expansion: run the vertex phase only; the frontier empties.
      tb("prologue");
      hipLaunchKernelGGL(k_frontier_prologue<ProgramOp<Program>>, dim3(grid_for(nf_host, 256, 4096)), dim3(256), 0, s,
                         ProgramOp<Program>{st}, f_in, &c->gather_size);
      te("prologue");
    }
    for (int p = 0; p < np; p++) {
      st.in_edges = passes[p];
      // The LBS expansion only marks the next frontier in dedup_next
      // (k_gas_order lists it); the tiered one appends it.
      GasState<Program> st_lbs = st;
      st_lbs.append = 0;
      st.append = 1;
      ProgramOp<Program> op_lbs{st_lbs};
      ProgramOp<Program> op{st};
      const bool first = p == 0;
      if (mode != 2) {
        static const char *names[4] = {"lbs_count", "lbs_scan", "lbs_partition", "lbs_expand"};
        LbsBuffers b{ws.scan_local, ws.partials, ws.info, ws.parts, ws.counters, config.expand_grid,
                     config.expand_vt};
        lbs_launch(op_lbs, first, f_in, &c->lbs_size, nf_host, max_path, b, f_out, sz_out, s, [&](int phase) {
          if (phase > 0) te(names[phase - 1]);
          if (phase < 4) tb(names[phase]);
        });
      }
      if (mode != 1) {
        const int grid = mode == 2 ? grid_for(nf_host, kExpandBlock) : tgrid;
        if (first && (st.do_apply || st.do_post)) {
          tb("prologue");
          hipLaunchKernelGGL(k_frontier_prologue<ProgramOp<Program>>, dim3(mode == 2 ? grid_for(nf_host, 256) : tgrid),
                             dim3(256), 0, s, op, f_in, &c->tier_size);
          te("prologue");
          tb("tiered_expand");
          hipLaunchKernelGGL((k_tiered_expand<ProgramOp<Program>, false>), dim3(grid), dim3(kExpandBlock), 0, s, op,
                             f_in, &c->tier_size, f_out, sz_out, config.wave_degree, config.block_degree,
                             ws.counters);
          te("tiered_expand");
        } else {
          // No apply phase: the prologue only clears dedup_prev, safe to fuse.
          tb("tiered_expand");
          if (first)
            hipLaunchKernelGGL((k_tiered_expand<ProgramOp<Program>, true>), dim3(grid), dim3(kExpandBlock), 0, s, op,
                               f_in, &c->tier_size, f_out, sz_out, config.wave_degree, config.block_degree,
                               ws.counters);
          else
            hipLaunchKernelGGL((k_tiered_expand<ProgramOp<Program>, false>), dim3(grid), dim3(kExpandBlock), 0, s, op,
                               f_in, &c->tier_size, f_out, sz_out, config.wave_degree, config.block_degree,
                               ws.counters);
          te("tiered_expand");
        }
      }
    }
    if (np > 0) {
      const int32_t words = (n + 31) / 32;
      const int grid = grid_for(words, kOrderBlock, kOrderGridMax);
      const int32_t per_thread = (int32_t)(((int64_t)words + (int64_t)grid * kOrderBlock - 1) / ((int64_t)grid * kOrderBlock));
      tb("order");
      hipLaunchKernelGGL(k_gas_order<>, dim3(grid), dim3(kOrderBlock), 0, s, c, st.dedup_next, words,
                         per_thread < 1 ? 1 : per_thread, sz_out, std::max<int32_t>(n / 64, 4096), f_out,
                         ws.order_words);
      te("order");
    }
  }

  // Read-back block: the loop state, both frontier sizes and the edge counter.
  struct Snapshot {
    GasCtl c;
    int32_t sizes[2];
    unsigned long long edges;
  };

  hipError_t snapshot(Workspace &ws, hipStream_t s, Snapshot *out) {
    char *h = reinterpret_cast<char *>(ws.h_ctl);
    GAS_RETURN_IF(hipMemcpyAsync(h, ws.ctl, sizeof(GasCtl), hipMemcpyDeviceToHost, s));
    GAS_RETURN_IF(hipMemcpyAsync(h + sizeof(GasCtl), ws.sizes, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    GAS_RETURN_IF(
        hipMemcpyAsync(h + sizeof(GasCtl) + 8, ws.counters, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    GAS_RETURN_IF(hipStreamSynchronize(s));
    std::memcpy(&out->c, h, sizeof(GasCtl));
    std::memcpy(out->sizes, h + sizeof(GasCtl), 8);
    std::memcpy(&out->edges, h + sizeof(GasCtl) + 8, 8);
    return hipSuccess;
  }

  hipError_t host_loop(CsrProblemT &problem, GasState<Program> st, int threshold, int iter_num) {
    hipStream_t s = problem.stream;
    Workspace &ws = problem.ws;
    Snapshot snap;
    GAS_RETURN_IF(snapshot(ws, s, &snap));
    int32_t nf = snap.sizes[0];
    for (int it = 0; it < iter_num && nf > 0; it++) {
      const int sel = it & 1;
      launch_gas_ctl(problem, sel, s);  // its choice of expansion is read back
      GAS_RETURN_IF(hipGetLastError());
      GAS_RETURN_IF(snapshot(ws, s, &snap));
      const bool lbs = snap.c.lbs_size > 0;
      issue_iteration(problem, st, sel, lbs ? 1 : 2, nf, 0, s, false);
      GAS_RETURN_IF(hipGetLastError());
      GAS_RETURN_IF(snapshot(ws, s, &snap));
      const int32_t next = snap.sizes[sel ^ 1];
      history.push_back({nf, (int64_t)snap.edges - total_edges, next, lbs ? 1 : 0});
      total_edges = (int64_t)snap.edges;
      if (verbose) std::printf("iteration %d: frontier %d -> %d (%s)\n", it, nf, next, lbs ? "lbs" : "tiered");
      nf = next;
    }
    iterations_run = (int32_t)history.size();
    for (auto &r : history) {
      frontier_total += r.frontier;
      next_total += r.next;
      lbs_iterations += r.lbs;
    }
    return hipSuccess;
  }

  hipError_t device_loop(CsrProblemT &problem, GasState<Program> st, int threshold) {
    hipStream_t s = problem.stream;
    Workspace &ws = problem.ws;
    const int32_t n = problem.graph.nodes;
    const int tgrid = tier_grid(threshold, n);
    std::vector<uint64_t> key;
    auto add = [&](const void *p, size_t bytes) {
      const unsigned char *b = static_cast<const unsigned char *>(p);
      for (size_t i = 0; i < bytes; i += 8) {
        uint64_t w = 0;
        std::memcpy(&w, b + i, std::min<size_t>(8, bytes - i));
        key.push_back(w);
      }
    };
    add(&st, sizeof(st));
    add(&config, sizeof(config));
    const void *bufs[] = {ws.frontier[0], ws.frontier[1], ws.dedup[0], ws.dedup[1], ws.sizes, ws.counters,
                          ws.scan_local, ws.partials, ws.parts, ws.info, ws.gcarry};
    add(bufs, sizeof(bufs));
    key.push_back((uint64_t)tgrid);
    if (!ws.loop_exec || key != ws.loop_key) {
      ws.ResetLoop();
      if (!ws.cap_stream) GAS_RETURN_IF(hipStreamCreateWithFlags(&ws.cap_stream, hipStreamNonBlocking));
      hipStream_t cs = ws.cap_stream;
      hipGraph_t graph = nullptr;
      GAS_RETURN_IF(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
      for (int j = 0; j < kLoopIters; j++) issue_iteration(problem, st, j & 1, 0, n, tgrid, cs);
      hipError_t e = hipStreamEndCapture(cs, &graph);
      if (e != hipSuccess) return e;
      e = hipGraphInstantiate(&ws.loop_exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (e != hipSuccess) {
        ws.loop_exec = nullptr;
        return e;
      }
      ws.loop_key = key;
    }
    Snapshot snap;
    for (;;) {
      GAS_RETURN_IF(hipGraphLaunch(ws.loop_exec, s));
      replays++;
      GAS_RETURN_IF(snapshot(ws, s, &snap));
      // Finished: the controller saw the end, or the replay's last iteration
      // left nothing (or reached iter_num) and no controller has run since.
      if (snap.c.done || snap.sizes[0] == 0 || snap.c.iterations >= snap.c.iter_num) break;
    }
    if (!snap.c.done) snap.c.sum_next += snap.sizes[0];
    iterations_run = snap.c.iterations;
    lbs_iterations = snap.c.lbs_iters;
    frontier_total = snap.c.sum_frontier;
    next_total = snap.c.sum_next;
    total_edges = (int64_t)snap.edges;
    return hipSuccess;
  }

  void tb(const char *name) {
    if (timer) timer->begin(name);
  }
  void te(const char *name) {
    if (timer) timer->end(name);
  }
};

}  // namespace GASengine
