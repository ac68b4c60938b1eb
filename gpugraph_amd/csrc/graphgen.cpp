// Host graph construction: Matrix Market reader and seeded generators
// (Erdos-Renyi G(n,M), R-MAT, percolated lattice) producing int32 CSR.
//
// Role in the reference: b40c/graph/builder/market.cuh:42-209 (MTX -> COO),
// b40c/graph/csr_graph.cuh:121-150 (FromCoo: stable sort by (row,col), row
// offsets), b40c/graph/builder/rmat.cuh (R-MAT, reseeded from time(NULL) on
// every draw there, so not reproducible).  Here every generator draws from a
// counter-based stream so the output depends only on (parameters, seed).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <unordered_set>
#include <vector>

#include "common.h"

namespace {

int num_threads() {
  const char *env = std::getenv("GG_HOST_THREADS");
  int t = env ? std::atoi(env) : 0;
  if (t <= 0) t = (int)std::thread::hardware_concurrency();
  if (t <= 0) t = 1;
  return std::min(t, 16);
}

// Run fn(lo, hi) over [0, count) split into contiguous chunks.
template <class F>
void parallel_for(int64_t count, F fn) {
  int t = num_threads();
  if (count < 65536 || t == 1) {
    fn((int64_t)0, count);
    return;
  }
  std::vector<std::thread> pool;
  int64_t chunk = (count + t - 1) / t;
  for (int i = 0; i < t; i++) {
    int64_t lo = i * chunk, hi = std::min(count, lo + chunk);
    if (lo >= hi) break;
    pool.emplace_back([=] { fn(lo, hi); });
  }
  for (auto &th : pool) th.join();
}

bool alloc_csr(gg_csr *out, int32_t n, int64_t m) {
  out->n = n;
  out->m = m;
  out->values = nullptr;
  out->value_dtype = GG_VALUES_NONE;
  out->row_offsets = (int32_t *)std::malloc(sizeof(int32_t) * ((size_t)n + 1));
  out->col_indices = (int32_t *)std::malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(m, 1));
  if (!out->row_offsets || !out->col_indices) {
    std::free(out->row_offsets);
    std::free(out->col_indices);
    out->row_offsets = out->col_indices = nullptr;
    return false;
  }
  return true;
}

// Build a CSR from directed (u, v) pairs: bucket by row, sort each row,
// optionally drop duplicates and self-loops.  Row contents are sorted, so the
// result is independent of the bucketing order and thread count.
int csr_from_pairs(int32_t n, const std::vector<uint64_t> &pairs, bool dedup,
                   bool drop_self, gg_csr *out) {
  const int64_t P = (int64_t)pairs.size();
  std::vector<std::atomic<int32_t>> deg((size_t)n);
  for (auto &d : deg) d.store(0, std::memory_order_relaxed);
  parallel_for(P, [&](int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; i++) {
      uint32_t u = (uint32_t)(pairs[i] >> 32), v = (uint32_t)pairs[i];
      if (drop_self && u == v) continue;
      deg[u].fetch_add(1, std::memory_order_relaxed);
    }
  });
  std::vector<int64_t> off((size_t)n + 1);
  off[0] = 0;
  for (int32_t v = 0; v < n; v++) off[v + 1] = off[v] + deg[v].load(std::memory_order_relaxed);
  const int64_t raw = off[n];
  std::vector<int32_t> tmp((size_t)std::max<int64_t>(raw, 1));
  std::vector<std::atomic<int64_t>> cur((size_t)n);
  for (int32_t v = 0; v < n; v++) cur[v].store(off[v], std::memory_order_relaxed);
  parallel_for(P, [&](int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; i++) {
      uint32_t u = (uint32_t)(pairs[i] >> 32), v = (uint32_t)pairs[i];
      if (drop_self && u == v) continue;
      tmp[(size_t)cur[u].fetch_add(1, std::memory_order_relaxed)] = (int32_t)v;
    }
  });
  std::vector<int32_t> newdeg((size_t)n);
  parallel_for(n, [&](int64_t lo, int64_t hi) {
    for (int64_t v = lo; v < hi; v++) {
      int32_t *b = tmp.data() + off[v], *e = tmp.data() + off[v + 1];
      std::sort(b, e);
      if (dedup) e = std::unique(b, e);
      newdeg[v] = (int32_t)(e - b);
    }
  });
  int64_t m = 0;
  for (int32_t v = 0; v < n; v++) m += newdeg[v];
  if (m >= INT32_MAX) return gg::set_error(GG_ERR_OVERFLOW, "graph has %lld edges; int32 offsets need m < 2^31", (long long)m);
  if (!alloc_csr(out, n, m)) return gg::set_error(GG_ERR_OOM, "host allocation of CSR (n=%d, m=%lld) failed", n, (long long)m);
  out->row_offsets[0] = 0;
  for (int32_t v = 0; v < n; v++) out->row_offsets[v + 1] = out->row_offsets[v] + newdeg[v];
  parallel_for(n, [&](int64_t lo, int64_t hi) {
    for (int64_t v = lo; v < hi; v++)
      std::memcpy(out->col_indices + out->row_offsets[v], tmp.data() + off[v], sizeof(int32_t) * (size_t)newdeg[v]);
  });
  return GG_OK;
}

// Seeded bijection on [0, 2^bits): odd multiply, add, xorshift rounds.
struct Scramble {
  uint64_t mask, mul[3], add[3];
  int shift;
  Scramble(int bits, uint64_t seed) {
    mask = bits >= 64 ? ~0ull : ((1ull << bits) - 1);
    shift = std::max(1, bits / 2);
    for (int r = 0; r < 3; r++) {
      mul[r] = gg::counter_rng(seed, 91, 2 * r) | 1ull;
      add[r] = gg::counter_rng(seed, 91, 2 * r + 1);
    }
  }
  uint64_t operator()(uint64_t x) const {
    for (int r = 0; r < 3; r++) {
      x = (x * mul[r] + add[r]) & mask;
      x ^= x >> shift;
    }
    return x;
  }
};

inline uint64_t pack(uint32_t u, uint32_t v) { return ((uint64_t)u << 32) | v; }

}  // namespace

extern "C" {

void gg_csr_free(gg_csr *csr) {
  if (!csr) return;
  std::free(csr->row_offsets);
  std::free(csr->col_indices);
  std::free(csr->values);
  csr->row_offsets = csr->col_indices = nullptr;
  csr->values = nullptr;
  csr->n = 0;
  csr->m = 0;
}

int gg_gen_erdos_renyi(int32_t n, int64_t num_pairs, uint64_t seed, gg_csr *out) {
  gg::clear_error();
  GG_CHECK_ARG(out, "out is NULL");
  GG_CHECK_ARG(n >= 2, "n must be >= 2 (got %d)", n);
  GG_CHECK_ARG(num_pairs >= 0 && num_pairs <= (int64_t)n * (n - 1) / 2, "num_pairs %lld out of range", (long long)num_pairs);
  std::unordered_set<uint64_t> seen;
  seen.reserve((size_t)num_pairs * 2);
  std::vector<uint64_t> pairs;
  pairs.reserve((size_t)num_pairs * 2);
  for (uint64_t i = 0; (int64_t)seen.size() < num_pairs; i++) {
    uint64_t r = gg::counter_rng(seed, 1, i);
    uint32_t u = (uint32_t)((r & 0xffffffffu) % (uint64_t)n);
    uint32_t v = (uint32_t)((r >> 32) % (uint64_t)n);
    if (u == v) continue;
    uint64_t key = pack(std::min(u, v), std::max(u, v));
    if (!seen.insert(key).second) continue;
    pairs.push_back(pack(u, v));
    pairs.push_back(pack(v, u));
  }
  return csr_from_pairs(n, pairs, true, true, out);
}

int gg_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                int32_t undirected, uint64_t seed, gg_csr *out) {
  gg::clear_error();
  GG_CHECK_ARG(out, "out is NULL");
  GG_CHECK_ARG(scale >= 1 && scale <= 30, "scale %d out of range [1,30]", scale);
  GG_CHECK_ARG(edge_factor >= 1, "edge_factor must be >= 1");
  GG_CHECK_ARG(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0, "invalid R-MAT probabilities");
  const int32_t n = (int32_t)(1u << scale);
  const int64_t draws = (int64_t)edge_factor << scale;
  GG_CHECK_ARG(draws * (undirected ? 2 : 1) < INT32_MAX, "too many edges for int32 offsets");
  const double two32 = 4294967296.0;
  const uint64_t ta = (uint64_t)(a * two32), tab = (uint64_t)((a + b) * two32),
                 tabc = (uint64_t)((a + b + c) * two32);
  Scramble perm(scale, seed ^ 0x5bd1e995ull);
  std::vector<uint64_t> pairs((size_t)draws * (undirected ? 2 : 1));
  parallel_for(draws, [&](int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; i++) {
      uint32_t u = 0, v = 0;
      uint64_t r = 0;
      for (int lvl = 0; lvl < scale; lvl++) {
        if ((lvl & 1) == 0) r = gg::counter_rng(seed, 2, (uint64_t)i * 16 + (uint64_t)(lvl >> 1));
        uint64_t x = (lvl & 1) ? (r >> 32) : (r & 0xffffffffu);
        uint32_t bit = 1u << (scale - 1 - lvl);
        if (x < ta) {
        } else if (x < tab) {
          v |= bit;
        } else if (x < tabc) {
          u |= bit;
        } else {
          u |= bit;
          v |= bit;
        }
      }
      u = (uint32_t)perm(u);
      v = (uint32_t)perm(v);
      if (undirected) {
        pairs[(size_t)(2 * i)] = pack(u, v);
        pairs[(size_t)(2 * i + 1)] = pack(v, u);
      } else {
        pairs[(size_t)i] = pack(u, v);
      }
    }
  });
  return csr_from_pairs(n, pairs, true, true, out);
}

int gg_gen_lattice(int32_t side, double keep_p, double long_range_frac, uint64_t seed, gg_csr *out) {
  gg::clear_error();
  GG_CHECK_ARG(out, "out is NULL");
  GG_CHECK_ARG(side >= 2 && (int64_t)side * side < INT32_MAX, "side %d out of range", side);
  GG_CHECK_ARG(keep_p >= 0 && keep_p <= 1 && long_range_frac >= 0, "invalid probabilities");
  const int64_t n = (int64_t)side * side;
  const int64_t H = (int64_t)side * (side - 1);  // horizontal bonds, then as many vertical
  const uint64_t thresh = keep_p >= 1.0 ? ~0ull : (uint64_t)(keep_p * 18446744073709551616.0);
  std::vector<unsigned char> keep((size_t)(2 * H));
  parallel_for(2 * H, [&](int64_t lo, int64_t hi) {
    for (int64_t bnd = lo; bnd < hi; bnd++) keep[(size_t)bnd] = gg::counter_rng(seed, 3, (uint64_t)bnd) < thresh;
  });
  int64_t kept = 0;
  for (auto k : keep) kept += k;
  const int64_t L = (int64_t)std::floor(long_range_frac * (double)kept);
  std::vector<uint64_t> pairs;
  pairs.reserve((size_t)(2 * (kept + L)));
  for (int64_t bnd = 0; bnd < 2 * H; bnd++) {
    if (!keep[(size_t)bnd]) continue;
    uint32_t u, v;
    if (bnd < H) {
      int64_t r = bnd / (side - 1), cc = bnd % (side - 1);
      u = (uint32_t)(r * side + cc);
      v = u + 1;
    } else {
      int64_t k = bnd - H;  // k = r*side + c for r < side-1
      u = (uint32_t)k;
      v = (uint32_t)(k + side);
    }
    pairs.push_back(pack(u, v));
    pairs.push_back(pack(v, u));
  }
  int64_t added = 0;
  for (uint64_t i = 0; added < L; i++) {
    uint64_t r = gg::counter_rng(seed, 4, i);
    uint32_t u = (uint32_t)((r & 0xffffffffu) % (uint64_t)n), v = (uint32_t)((r >> 32) % (uint64_t)n);
    if (u == v) continue;
    pairs.push_back(pack(u, v));
    pairs.push_back(pack(v, u));
    added++;
  }
  return csr_from_pairs((int32_t)n, pairs, true, true, out);
}

int gg_csr_weights(gg_csr *csr, int32_t value_dtype, int32_t lo, int32_t hi, int32_t symmetric, uint64_t seed) {
  gg::clear_error();
  GG_CHECK_ARG(csr && csr->row_offsets && csr->col_indices, "csr is empty");
  GG_CHECK_ARG(value_dtype == GG_VALUES_I32 || value_dtype == GG_VALUES_F32, "value_dtype must be I32 or F32");
  GG_CHECK_ARG(value_dtype != GG_VALUES_I32 || hi >= lo, "hi < lo");
  void *vals = std::malloc(4 * (size_t)std::max<int64_t>(csr->m, 1));
  if (!vals) return gg::set_error(GG_ERR_OOM, "host allocation of %lld values failed", (long long)csr->m);
  const int32_t n = csr->n;
  const int32_t *ro = csr->row_offsets, *col = csr->col_indices;
  const uint64_t span = (uint64_t)((int64_t)hi - lo + 1);
  parallel_for(n, [&](int64_t a, int64_t b) {
    for (int64_t u = a; u < b; u++) {
      for (int32_t e = ro[u]; e < ro[u + 1]; e++) {
        uint32_t v = (uint32_t)col[e];
        uint64_t r;
        if (symmetric)
          r = gg::counter_rng(seed, 5, pack(std::min<uint32_t>((uint32_t)u, v), std::max<uint32_t>((uint32_t)u, v)));
        else
          r = gg::counter_rng(seed, 6, (uint64_t)e);
        if (value_dtype == GG_VALUES_I32)
          ((int32_t *)vals)[e] = lo + (int32_t)(r % span);
        else
          ((float *)vals)[e] = (float)((double)((r >> 40) + 1) / 16777216.0);
      }
    }
  });
  std::free(csr->values);
  csr->values = vals;
  csr->value_dtype = value_dtype;
  return GG_OK;
}

int gg_read_mtx(const char *path, int32_t undirected, int32_t value_dtype, gg_csr *out) {
  gg::clear_error();
  GG_CHECK_ARG(path && out, "path/out is NULL");
  GG_CHECK_ARG(value_dtype == GG_VALUES_NONE || value_dtype == GG_VALUES_I32 || value_dtype == GG_VALUES_F32,
               "bad value_dtype %d", value_dtype);
  FILE *f = std::fopen(path, "r");
  if (!f) return gg::set_error(GG_ERR_IO, "cannot open %s", path);
  struct Coo {
    int32_t row, col;
    double val;
  };
  std::vector<Coo> coo;
  char line[1024];
  int64_t nodes = -1, declared = 0;
  int rc = GG_OK;
  while (std::fgets(line, sizeof line, f)) {
    if (line[0] == '%') continue;
    char *p = line;
    while (*p == ' ' || *p == '\t') p++;
    if (*p == '\n' || *p == '\0') continue;
    if (nodes < 0) {
      long long nx, ny, ne;
      if (std::sscanf(p, "%lld %lld %lld", &nx, &ny, &ne) != 3) {
        rc = gg::set_error(GG_ERR_IO, "%s: invalid problem description", path);
        break;
      }
      if (nx != ny) {
        rc = gg::set_error(GG_ERR_IO, "%s: not square (%lld, %lld)", path, nx, ny);
        break;
      }
      if (nx >= INT32_MAX || ne * (undirected ? 2 : 1) >= INT32_MAX) {
        rc = gg::set_error(GG_ERR_OVERFLOW, "%s: too large for int32 ids", path);
        break;
      }
      nodes = nx;
      declared = ne;
      coo.reserve((size_t)(ne * (undirected ? 2 : 1)));
      continue;
    }
    long long r, c;
    double val = 1.0;
    int got = std::sscanf(p, "%lld %lld %lf", &r, &c, &val);
    if (got < 2) {
      rc = gg::set_error(GG_ERR_IO, "%s: badly formed edge line", path);
      break;
    }
    if ((int64_t)(coo.size() / (undirected ? 2 : 1)) >= declared) {
      rc = gg::set_error(GG_ERR_IO, "%s: more than %lld edges", path, (long long)declared);
      break;
    }
    if (r < 1 || c < 1 || r > nodes || c > nodes) {
      rc = gg::set_error(GG_ERR_IO, "%s: vertex id out of range", path);
      break;
    }
    coo.push_back({(int32_t)(r - 1), (int32_t)(c - 1), val});
    if (undirected) coo.push_back({(int32_t)(c - 1), (int32_t)(r - 1), val});
  }
  std::fclose(f);
  if (rc != GG_OK) return rc;
  if (nodes < 0) return gg::set_error(GG_ERR_IO, "%s: no graph found", path);
  if ((int64_t)coo.size() != declared * (undirected ? 2 : 1))
    return gg::set_error(GG_ERR_IO, "%s: only %lld/%lld edges read", path, (long long)coo.size(), (long long)declared);
  std::stable_sort(coo.begin(), coo.end(), [](const Coo &x, const Coo &y) {
    return x.row < y.row || (x.row == y.row && x.col < y.col);
  });
  const int64_t m = (int64_t)coo.size();
  if (!alloc_csr(out, (int32_t)nodes, m)) return gg::set_error(GG_ERR_OOM, "host allocation failed");
  if (value_dtype != GG_VALUES_NONE) {
    out->values = std::malloc(4 * (size_t)std::max<int64_t>(m, 1));
    if (!out->values) {
      gg_csr_free(out);
      return gg::set_error(GG_ERR_OOM, "host allocation failed");
    }
    out->value_dtype = value_dtype;
  }
  int64_t e = 0;
  for (int32_t v = 0; v <= (int32_t)nodes; v++) {
    while (e < m && coo[(size_t)e].row < v) e++;
    out->row_offsets[v] = (int32_t)e;
  }
  for (int64_t i = 0; i < m; i++) {
    out->col_indices[i] = coo[(size_t)i].col;
    if (value_dtype == GG_VALUES_I32) ((int32_t *)out->values)[i] = (int32_t)coo[(size_t)i].val;
    if (value_dtype == GG_VALUES_F32) ((float *)out->values)[i] = (float)coo[(size_t)i].val;
  }
  return GG_OK;
}

}  // extern "C"
