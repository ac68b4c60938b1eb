// Automatic tuner over gg_options (gg_autotune).
//
// The reference fixes its schedule knobs at compile time (degree tiers
// 32/512, enactor_vertex_centric.cuh:3949-3950) or takes them from the command
// line (threshold, Algorithms/SSSP/register.h:39).  Here every knob that only
// changes the schedule -- LBS tile shape, expand grid, degree tiers, the
// LBS/tiered switch, the direction-optimising BFS switch points, the SSSP
// bucket width and pull switch -- is searched on the caller's own graph:
// coordinate descent over a candidate table.  Each trial times the programs
// the knob affects with the library's HIP-event device time, the candidate's
// runs bracketed by runs of the current options (fastest of `reps` runs each,
// so clock drift and one-off stalls cancel), and is accepted when the
// candidate is at least 3% faster.  A final bracketed comparison against the
// starting options keeps the result only if it still wins.  Every trial's integer outputs
// are checksummed on the device and must equal the baseline's: a knob that
// changed a result would be a bug, and the tuner reports it as one.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>

#include "internal.h"

namespace gg {
namespace {

// sum over v of (val[v] + 1) * mix64(v), mod 2^64 (order-independent).
__global__ void k_checksum(const int32_t *__restrict__ val, int32_t n, unsigned long long *out) {
  unsigned long long acc = 0;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
    unsigned long long h = (unsigned long long)v * 0x9E3779B97F4A7C15ull;
    h ^= h >> 29;
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 32;
    acc += ((unsigned long long)(uint32_t)val[v] + 1ull) * h;
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if ((threadIdx.x & 63) == 0 && acc) atomicAdd(out, acc);
}

// sum over v of |a[v] - b[v]| in FP64.
__global__ void k_l1_diff(const float *__restrict__ a, const float *__restrict__ b, int32_t n, double *out) {
  double acc = 0.0;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x)
    acc += fabs((double)a[v] - (double)b[v]);
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if ((threadIdx.x & 63) == 0 && acc != 0.0) atomicAdd(out, acc);
}

constexpr double kPrTol = 1e-6, kPrAgree = 1e-5;

struct Program {
  uint32_t bit;
  int run(gg_graph *g, int32_t src, int32_t *out, gg_stats *st) const {
    switch (bit) {
      case GG_TUNE_BFS:
        return gg_bfs(g, src, out, GG_OUT_DEVICE, st);
      case GG_TUNE_SSSP:
        return gg_sssp(g, src, out, GG_OUT_DEVICE, st);
      case GG_TUNE_CC:
        return gg_cc(g, out, GG_OUT_DEVICE, st);
      case GG_TUNE_PAGERANK:
        return gg_pagerank(g, 0.85f, nullptr, kPrTol, 1000, reinterpret_cast<float *>(out), GG_OUT_DEVICE, st);
      default:
        return set_error(GG_ERR_INVALID, "unknown program bit %u", bit);
    }
  }
};

// GG_TUNE_GAS runs two reference-form programs: push-only BFS and GAS SSSP.
struct Runner {
  gg_graph *g;
  int32_t src;
  int32_t reps;
  int32_t *out;
  unsigned long long *sum;
  float *pr_ref;       // baseline PageRank result
  bool pr_ref_set;

  int checksum(unsigned long long *h) {
    GG_HIP(hipMemsetAsync(sum, 0, 8, g->stream));
    hipLaunchKernelGGL(k_checksum, dim3(GASengine::grid_for(g->n, 256, 1024)), dim3(256), 0, g->stream, out, g->n, sum);
    GG_HIP(hipGetLastError());
    GG_HIP(hipMemcpyAsync(h, sum, 8, hipMemcpyDeviceToHost, g->stream));
    GG_HIP(hipStreamSynchronize(g->stream));
    return GG_OK;
  }

  // PageRank results are FP32 sums whose rounding depends on the schedule:
  // the first run is kept as the reference, later ones must agree within
  // L1 kPrAgree; *h = 0 when they do, 1 when not.
  int pr_check(unsigned long long *h) {
    if (!pr_ref_set) {
      GG_HIP(hipMemcpyAsync(pr_ref, out, 4 * (size_t)g->n, hipMemcpyDeviceToDevice, g->stream));
      pr_ref_set = true;
      *h = 0;
      return GG_OK;
    }
    double *d = reinterpret_cast<double *>(sum);
    GG_HIP(hipMemsetAsync(d, 0, 8, g->stream));
    hipLaunchKernelGGL(k_l1_diff, dim3(GASengine::grid_for(g->n, 256, 1024)), dim3(256), 0, g->stream,
                       reinterpret_cast<const float *>(out), pr_ref, g->n, d);
    GG_HIP(hipGetLastError());
    double l1 = 0.0;
    GG_HIP(hipMemcpyAsync(&l1, d, 8, hipMemcpyDeviceToHost, g->stream));
    GG_HIP(hipStreamSynchronize(g->stream));
    *h = l1 <= kPrAgree ? 0 : 1;
    return GG_OK;
  }

  // One program class: fastest device ms over reps, checksum(s) of its outputs.
  int time_one(uint32_t bit, double *ms, unsigned long long *hsum) {
    std::vector<double> t;
    unsigned long long acc = 0;
    for (int r = 0; r < reps; r++) {
      double total = 0.0;
      unsigned long long hs = 0;
      if (bit == GG_TUNE_GAS) {
        gg_stats st{};
        GG_TRY(gg_bfs(g, src, out, GG_OUT_DEVICE | GG_BFS_PUSH_ONLY, &st));
        total += st.gpu_ms;
        unsigned long long h1 = 0;
        if (r == 0) GG_TRY(checksum(&h1));
        hs += h1;
        if (g->value_dtype == GG_VALUES_I32) {
          GG_TRY(gg_sssp(g, src, out, GG_OUT_DEVICE | GG_SSSP_GAS, &st));
          total += st.gpu_ms;
          unsigned long long h2 = 0;
          if (r == 0) GG_TRY(checksum(&h2));
          hs += 3 * h2;
        }
      } else {
        gg_stats st{};
        GG_TRY(Program{bit}.run(g, src, out, &st));
        total += st.gpu_ms;
        if (r == 0) GG_TRY(bit == GG_TUNE_PAGERANK ? pr_check(&hs) : checksum(&hs));
      }
      if (r == 0) acc = hs;
      t.push_back(total);
    }
    *ms = *std::min_element(t.begin(), t.end());
    *hsum = acc;
    return GG_OK;
  }
};

// Time options `a` and `b` on the programs in `hit`: a, b, a (reps runs each,
// fastest kept, a's two rounds merged).  b's outputs must match ref[].
int bracket(gg_graph *g, Runner &run, const gg_options &a, const gg_options &b, uint32_t hit, const uint32_t *bits,
            int nbits, const unsigned long long *ref, double *ms_a, double *ms_b, int *bad_prog) {
  *ms_a = *ms_b = 0.0;
  *bad_prog = 0;
  for (int p = 0; p < nbits; p++) {
    if (!(hit & bits[p])) continue;
    double ta1 = 0, tb = 0, ta2 = 0;
    unsigned long long h = 0;
    GG_TRY(gg_graph_set_options(g, &a));
    GG_TRY(run.time_one(bits[p], &ta1, &h));
    GG_TRY(gg_graph_set_options(g, &b));
    GG_TRY(run.time_one(bits[p], &tb, &h));
    if (h != ref[p]) *bad_prog = (int)bits[p];
    GG_TRY(gg_graph_set_options(g, &a));
    GG_TRY(run.time_one(bits[p], &ta2, &h));
    *ms_a += std::min(ta1, ta2);
    *ms_b += tb;
  }
  return GG_OK;
}

struct Knob {
  int field;           // index of the int32 field in gg_options
  uint32_t affects;    // GG_TUNE_* programs whose schedule it changes
  std::vector<int32_t> values;
};

int32_t &field_ref(gg_options &o, int field) { return reinterpret_cast<int32_t *>(&o)[field]; }

constexpr int kFieldThreshold = 0, kFieldWave = 1, kFieldBlock = 2, kFieldAlpha = 4, kFieldBeta = 5, kFieldTile = 6,
              kFieldDelta = 7, kFieldPull = 8, kFieldGrid = 9, kFieldSlices = 10, kFieldGasGrid = 11;
static_assert(sizeof(gg_options) == 12 * sizeof(int32_t), "gg_options layout");

std::vector<Knob> knob_table() {
  const uint32_t lbs_users = GG_TUNE_BFS | GG_TUNE_SSSP | GG_TUNE_CC | GG_TUNE_GAS;
  return {
      {kFieldTile, lbs_users, {512, 1024, 2048}},
      {kFieldGrid, GG_TUNE_BFS | GG_TUNE_SSSP | GG_TUNE_CC, {512, 1024, 2048, 4096, 8192}},
      {kFieldGasGrid, GG_TUNE_GAS, {2048, 4096, 8192, 16384}},
      {kFieldAlpha, GG_TUNE_BFS, {4, 8, 14, 24, 48}},
      {kFieldBeta, GG_TUNE_BFS, {8, 16, 24, 48, 96}},
      {kFieldWave, GG_TUNE_BFS | GG_TUNE_CC | GG_TUNE_GAS, {32, 64, 128, 256, 512}},
      {kFieldBlock, GG_TUNE_CC | GG_TUNE_GAS, {512, 1024, 2048, 4096, 16384}},
      {kFieldThreshold, GG_TUNE_CC | GG_TUNE_GAS, {0, 1000, 10000, 100000, INT_MAX}},
      // The pull threshold before the bucket width: small widths pay only
      // with a pull (without one, a run of narrow buckets never gathers a
      // frontier large enough to trigger it later either).  3 is tried
      // before 2: where both time alike (within the 3% a change must win
      // by), the search keeps m / 3, whose trigger sits further from the
      // frontier sizes of narrow buckets than m / 2 does.
      {kFieldPull, GG_TUNE_SSSP, {0, 3, 2, 4, 8, 16}},
      {kFieldDelta, GG_TUNE_SSSP, {3, 4, 6, 8, 12, 16, 24, 32, 64, 128, 256, 512, 1024}},
      {kFieldSlices, GG_TUNE_PAGERANK, {0, 1}},
  };
}

}  // namespace
}  // namespace gg

extern "C" int gg_autotune(gg_graph *g, uint32_t programs, int32_t src, int32_t reps, gg_options *best,
                           gg_tune_trial *trials, int32_t max_trials, int32_t *num_trials) {
  using namespace gg;
  clear_error();
  GG_CHECK_ARG(g, "graph is NULL");
  const uint32_t all = GG_TUNE_BFS | GG_TUNE_SSSP | GG_TUNE_CC | GG_TUNE_GAS | GG_TUNE_PAGERANK;
  GG_CHECK_ARG(programs && (programs & ~all) == 0, "programs must be a non-empty set of GG_TUNE_* bits");
  GG_CHECK_ARG(src >= 0 && src < g->n, "src %d out of range", src);
  GG_CHECK_ARG(reps >= 1 && reps <= 64, "reps must be in [1, 64]");
  GG_CHECK_ARG(max_trials >= 0 && (max_trials == 0 || trials), "trials is NULL");
  if (g->value_dtype != GG_VALUES_I32) programs &= ~GG_TUNE_SSSP;
  if (g->directed) programs &= ~GG_TUNE_CC;  // CC is defined on undirected graphs
  GG_CHECK_ARG(programs, "no selected program runs on this graph (SSSP needs int32 values, CC an undirected graph)");
  GG_TRY(activate(g));

  DevBuf outb, sumb, refb;
  GG_TRY(outb.ensure(4 * (size_t)g->n));
  GG_TRY(sumb.ensure(8));
  if (programs & GG_TUNE_PAGERANK) GG_TRY(refb.ensure(4 * (size_t)g->n));
  Runner run{g, src, reps, outb.as<int32_t>(), sumb.as<unsigned long long>(), refb.as<float>(), false};
  constexpr int kProgs = 5;
  const uint32_t bits[kProgs] = {GG_TUNE_BFS, GG_TUNE_SSSP, GG_TUNE_CC, GG_TUNE_GAS, GG_TUNE_PAGERANK};

  const gg_options start = g->opt;
  gg_options cur = start;
  double cur_ms[kProgs] = {};  // baseline timings (logged)
  unsigned long long ref_sum[kProgs] = {};
  int32_t nt = 0;
  auto log_trial = [&](uint32_t progs, int field, int32_t value, double ms) {
    if (nt < max_trials) trials[nt] = gg_tune_trial{progs, field, value, (float)ms};
    nt++;
  };
  auto restore = [&](int rc) {
    g->opt = start;
    return rc;
  };
  // Baseline.
  double base_total = 0.0;
  for (int p = 0; p < kProgs; p++) {
    if (!(programs & bits[p])) continue;
    int rc = run.time_one(bits[p], &cur_ms[p], &ref_sum[p]);
    if (rc != GG_OK) return restore(rc);
    base_total += cur_ms[p];
  }
  log_trial(programs, -1, 0, base_total);

  const std::vector<Knob> knobs = knob_table();
  for (int pass = 0; pass < 2; pass++) {
    bool changed = false;
    for (const Knob &k : knobs) {
      const uint32_t hit = k.affects & programs;
      if (!hit) continue;
      for (int32_t val : k.values) {
        if (val == field_ref(cur, k.field)) continue;
        gg_options trial = cur;
        field_ref(trial, k.field) = val;
        if (trial.block_degree < trial.wave_degree) continue;
        double old_ms = 0.0, new_ms = 0.0;
        int bad = 0;
        int rc = bracket(g, run, cur, trial, hit, bits, kProgs, ref_sum, &old_ms, &new_ms, &bad);
        if (rc != GG_OK) return restore(rc);
        if (bad) {
          restore(GG_OK);
          return set_error(GG_ERR_INTERNAL, "autotune: option %d = %d changed the result of program %d", k.field, val,
                           bad);
        }
        log_trial(hit, k.field, val, new_ms);
        if (new_ms < 0.97 * old_ms) {
          cur = trial;
          changed = true;
        }
      }
    }
    if (!changed) break;
  }
  // Keep the result only if it still beats the starting options.
  if (std::memcmp(&cur, &start, sizeof(cur)) != 0) {
    double start_ms = 0.0, cur_ms_total = 0.0;
    int bad = 0;
    int rc = bracket(g, run, start, cur, programs, bits, kProgs, ref_sum, &start_ms, &cur_ms_total, &bad);
    if (rc != GG_OK) return restore(rc);
    log_trial(programs, -2, 0, cur_ms_total);
    if (!(cur_ms_total < start_ms)) cur = start;
  }
  g->opt = cur;
  if (best) *best = cur;
  if (num_trials) *num_trials = std::min(nt, max_trials);
  return GG_OK;
}
