// gg_bfs_noalloc: BFS over a CSR the caller already holds on the device,
// allocating nothing.  The frontier queues, visited bitmap, LBS scan and
// partition arrays and counters live in `scratch`, a fixed device array of
// this library (kNoAllocScratchBytes, about 25 MB, held on every GPU the
// library runs on until the process ends), so graphs are limited to
// kNoAllocMaxN vertices and kNoAllocMaxM edges.  Same result as gg_bfs
// (the reference BFS program's depths, Algorithms/BFS/bfs.h): top-down LBS
// expansion per level with the shared LBS kernels (gasengine/lbs.h) and a
// test-and-set in the visited bitmap; one 4-byte read-back per level.
//
// Every call on a GPU uses that GPU's one scratch array: calls that overlap
// on one GPU (two host threads, or two streams) race and give wrong results.
#include <hip/hip_runtime.h>

#include <mutex>

#include "internal.h"

namespace gg {
namespace {

using namespace GASengine;

constexpr int32_t kNoAllocMaxN = 1 << 21;
constexpr int64_t kNoAllocMaxM = (int64_t)1 << 26;

constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
// Scratch layout (byte offsets, each region 256-byte aligned).
constexpr size_t kOffQueue0 = 0;
constexpr size_t kOffQueue1 = kOffQueue0 + align256(4 * (size_t)kNoAllocMaxN);
constexpr size_t kOffScan = kOffQueue1 + align256(4 * (size_t)kNoAllocMaxN);
constexpr size_t kOffPartials = kOffScan + align256(4 * (size_t)kNoAllocMaxN);
constexpr size_t kOffParts = kOffPartials + align256(4 * ((size_t)kNoAllocMaxN / kCountTile + 2));
constexpr size_t kOffBits = kOffParts + align256(4 * ((size_t)(kNoAllocMaxN + kNoAllocMaxM) / kLbsTileMin + 2));
constexpr size_t kOffInfo = kOffBits + align256(4 * (((size_t)kNoAllocMaxN + 31) / 32));
constexpr size_t kOffSizes = kOffInfo + align256(sizeof(LbsInfo));
constexpr size_t kOffCounters = kOffSizes + 256;
constexpr size_t kNoAllocScratchBytes = kOffCounters + align256(8 * (size_t)kCounterWords);

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// Device address of `scratch` per device (hipGetSymbolAddress once per
// device, with that device current).  The array's name in host code is not a
// device address and is never passed to a kernel.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kMaxDevices] = {};

int scratch_of_current_device(unsigned char **out) {
  int dev = 0;
  GG_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return set_error(GG_ERR_UNSUPPORTED, "device %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(GG_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<unsigned char *>(p);
  }
  *out = g_scratch_ptr[dev];
  return GG_OK;
}

// Top-down edge operation: first touch wins the visited bit, stamps the
// depth and enqueues the vertex; mf = out-degree sum of the new frontier.
struct NoAllocPushOp {
  static constexpr bool kItemValues = false;
  const int32_t *ro;
  const int32_t *col;
  int32_t *depth;
  uint32_t *visited;
  unsigned long long *counters;
  int32_t next_depth;
  unsigned long long mf;

  __device__ __forceinline__ void prologue(int32_t) {}
  __device__ __forceinline__ int32_t degree(int32_t v) { return ro[v + 1] - ro[v]; }
  __device__ __forceinline__ int32_t row_begin(int32_t v) { return ro[v]; }
  __device__ __forceinline__ int32_t item_value(int32_t) { return 0; }
  __device__ __forceinline__ bool edge(int32_t, int32_t, int32_t slot, int32_t &out) {
    const int32_t nb = col[slot];
    if (!bit_test_and_set(visited, nb)) return false;
    depth[nb] = next_depth;
    mf += (unsigned long long)(ro[nb + 1] - ro[nb]);
    out = nb;
    return true;
  }
  __device__ __forceinline__ void finish() { block_counter_add<kExpandBlock>(&counters[1], mf); }
};

// Start state: depth -1 except src, visited = {src}, queue 0 = [src],
// sizes = {1, 0}, counters 0.
__global__ void __launch_bounds__(256)
k_noalloc_start(int32_t *__restrict__ depth, uint32_t *__restrict__ visited, int32_t n, int32_t src,
                int32_t *__restrict__ queue0, int32_t *__restrict__ sizes, unsigned long long *__restrict__ counters,
                const int32_t *__restrict__ ro) {
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
  const int32_t words = (n + 31) / 32;
  for (int64_t i = tid; i < n; i += stride) depth[i] = i == src ? 0 : -1;
  for (int64_t w = tid; w < words; w += stride) visited[w] = w == (src >> 5) ? 1u << (src & 31) : 0u;
  for (int64_t c = tid; c < kCounterWords; c += stride) counters[c] = 0ull;
  if (tid == 0) {
    queue0[0] = src;
    sizes[0] = 1;
    sizes[1] = 0;
    counters[1] = (unsigned long long)(ro[src + 1] - ro[src]);  // the source's out-degree joins the reached edges
  }
}

__global__ void k_noalloc_zero(int32_t *p) { *p = 0; }

}  // namespace
}  // namespace gg

extern "C" int gg_bfs_noalloc(int32_t n, int64_t m, const int32_t *d_row_offsets, const int32_t *d_col_indices,
                              int32_t src, int32_t *d_depth, void *stream, gg_stats *stats) {
  using namespace gg;
  gg::clear_error();
  GG_CHECK_ARG(d_row_offsets && d_col_indices && d_depth, "NULL device array");
  GG_CHECK_ARG(n >= 1 && m >= 0, "need n >= 1 and m >= 0");
  GG_CHECK_ARG(n <= kNoAllocMaxN, "n = %d exceeds the allocation-free BFS limit of %d vertices", n, kNoAllocMaxN);
  GG_CHECK_ARG(m <= kNoAllocMaxM, "m = %lld exceeds the allocation-free BFS limit of %lld edges", (long long)m,
               (long long)kNoAllocMaxM);
  GG_CHECK_ARG(src >= 0 && src < n, "src %d out of range [0, %d)", src, n);
  hipStream_t s = static_cast<hipStream_t>(stream);
  unsigned char *base = nullptr;
  GG_TRY(scratch_of_current_device(&base));
  int32_t *queue[2] = {reinterpret_cast<int32_t *>(base + kOffQueue0), reinterpret_cast<int32_t *>(base + kOffQueue1)};
  int32_t *sizes = reinterpret_cast<int32_t *>(base + kOffSizes);
  unsigned long long *ctr = reinterpret_cast<unsigned long long *>(base + kOffCounters);
  uint32_t *visited = reinterpret_cast<uint32_t *>(base + kOffBits);
  LbsBuffers lb{reinterpret_cast<int32_t *>(base + kOffScan),
                reinterpret_cast<int32_t *>(base + kOffPartials),
                reinterpret_cast<LbsInfo *>(base + kOffInfo),
                reinterpret_cast<int32_t *>(base + kOffParts),
                ctr + 2,
                2048,
                kExpandVt};
  hipLaunchKernelGGL(k_noalloc_start, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, d_depth, visited, n, src,
                     queue[0], sizes, ctr, d_row_offsets);
  GG_HIP(hipGetLastError());
  NoAllocPushOp op{d_row_offsets, d_col_indices, d_depth, visited, ctr, 0, 0ull};
  int32_t nf = 1, level = 0;
  int64_t reached = 1;
  while (nf > 0) {
    const int cur = level & 1, nxt = cur ^ 1;
    hipLaunchKernelGGL(k_noalloc_zero, dim3(1), dim3(1), 0, s, sizes + nxt);
    op.next_depth = level + 1;
    lbs_launch(op, false, queue[cur], sizes + cur, nf, (int64_t)n + m, lb, queue[nxt], sizes + nxt, s, [](int) {});
    GG_HIP(hipGetLastError());
    int32_t h = 0;
    GG_HIP(hipMemcpyAsync(&h, sizes + nxt, sizeof(h), hipMemcpyDeviceToHost, s));
    GG_HIP(hipStreamSynchronize(s));
    nf = h;
    reached += nf;
    level++;
  }
  if (stats) {
    unsigned long long hc[kCounterWords];
    GG_HIP(hipMemcpyAsync(hc, ctr, sizeof(hc), hipMemcpyDeviceToHost, s));
    GG_HIP(hipStreamSynchronize(s));
    unsigned long long c[3] = {0, 0, 0};
    for (int q = 0; q < kStripes; q++)
      for (int i = 0; i < 3; i++) c[i] += hc[(size_t)q * kCounterSlots + i];
    std::memset(stats, 0, sizeof(*stats));
    stats->iterations = level;
    stats->reached_vertices = reached;
    stats->reached_edges = (int64_t)c[1];
    stats->edges_inspected = (int64_t)c[2];
    stats->alg_bytes = 12.0 * (double)reached + 8.0 * (double)c[2] + 8.0 * (double)reached;
  }
  return GG_OK;
}
