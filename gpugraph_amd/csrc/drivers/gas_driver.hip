// Command-line drivers over the header-only vertex-program API, one binary
// per program (built with -DGG_PROGRAM_BFS, _SSSP, _CC or _PAGERANK).  The
// sequence is the reference drivers' (Algorithms/SSSP/sssp.cu:199-430,
// BFS/bfs.cu:845-881, CC/cc.cu:255-316, PageRank/PageRank.cu:255-326) with
// the HIP runtime in place of CUDA's: Config from -p / -c, BuildMarketGraph
// into a CsrGraph, CsrProblem(cfg) + the 12-argument FromHostProblem,
// EnactorVertexCentric(cfg, verbose), one EnactIterativeSearch per source,
// ExtractResults (and, for BFS, ExtractPreds).
//
//   -g / -graph FILE       Matrix Market graph (1-based ids)
//   -o / -output FILE      the result, one line per vertex (%d, or %.9g for
//                          PageRank ranks)
//   -s / -sources FILE     starting vertices, one per line (`origin` applies);
//                          run one after another, the output is the last one's
//   -p / -parameters STR   "name=value ..." (the parameters registered below)
//   -c FILE                parameter lines from a file
//   -preds FILE            BFS only: the BFS-tree predecessor of every vertex
//                          (the source its own, -1 where unreached)
//   -help
//
// Exit status 0 on success, 1 on any error (unknown parameter, unreadable
// graph, source out of range, a failed HIP call), with a message on stdout.
#include <climits>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "gasengine/config.h"
#include "gasengine/enactor_vertex_centric.h"
#include "gasengine/market.h"

#if defined(GG_PROGRAM_BFS)
#include "programs/bfs.h"
typedef bfs Program;
static const char *kProgram = "BFS";
static constexpr bool kWithValue = false;
static constexpr bool kMarkPredecessors = true;
#elif defined(GG_PROGRAM_SSSP)
#include "programs/sssp.h"
typedef sssp Program;
static const char *kProgram = "SSSP";
static constexpr bool kWithValue = true;
static constexpr bool kMarkPredecessors = false;
#elif defined(GG_PROGRAM_CC)
#include "programs/cc.h"
typedef cc Program;
static const char *kProgram = "CC";
static constexpr bool kWithValue = false;
static constexpr bool kMarkPredecessors = false;
#elif defined(GG_PROGRAM_PAGERANK)
#include "programs/pagerank.h"
typedef pagerank Program;
static const char *kProgram = "PageRank";
static constexpr bool kWithValue = false;
static constexpr bool kMarkPredecessors = false;
#else
#error "build with one of -DGG_PROGRAM_BFS, -DGG_PROGRAM_SSSP, -DGG_PROGRAM_CC, -DGG_PROGRAM_PAGERANK"
#endif

using namespace b40c;
using namespace graph;

typedef int VertexId;
typedef int SizeT;
typedef Program::DataType Value;

// The drivers' parameters (the union of the reference's register.h files).
void registerParameters() {
  Config::registerParameter<int>("src", "Starting vertex (default 1)", 1);
  Config::registerParameter<int>("verbose", "Print the frontier size of every iteration (default 0)", 0);
  Config::registerParameter<int>("origin", "Origin (0 or 1) of the vertex ids given (default 1)", 1);
  Config::registerParameter<int>("directed", "The graph is directed (default 1)", 1);
  Config::registerParameter<int>("device", "The device to use (default 0)", 0);
  Config::registerParameter<int>("iter_num", "Iterations to run at most (default INT_MAX)", INT_MAX);
  Config::registerParameter<int>("threshold",
                                 "Frontiers larger than this expand by load-balanced search, smaller ones by the "
                                 "tiered kernel (default 10000)",
                                 10000);
  Config::registerParameter<int>("with_value", "Load edge values from the Matrix Market file (default 1)", 1);
  Config::registerParameter<int>("num_src", "Starting vertices for a random source set (default 1)", 1);
  Config::registerParameter<int>("run_CPU", "Accepted for the reference's scripts; checks run in the tests", 0);
  Config::registerParameter<double>("max_queue_sizing", "Accepted for the reference's scripts; queues hold n", 1.5);
  Config::registerParameter<double>("tol", "Accepted for the reference's scripts; PageRank stops at |delta| < 0.01",
                                    0.01);
}

static void usage(const char *exe) {
  std::printf("%s driver of the GAS engine\n", kProgram);
  std::printf("Usage: %s -g graph.mtx [-o out.txt] [-s sources.txt] [-p \"name=value ...\"] [-c config.txt]%s\n",
              exe, kMarkPredecessors ? " [-preds preds.txt]" : "");
  std::printf("     -help display the command options\n");
  std::printf("     -graph or -g: a sparse matrix in Matrix Market (.mtx) format\n");
  std::printf("     -output or -o: file for the result\n");
  std::printf("     -sources or -s: file of starting vertices\n");
  std::printf("     -c: parameter lines from a file\n");
  if (kMarkPredecessors) std::printf("     -preds: file for the BFS-tree predecessors\n");
  std::printf("     -parameters or -p: parameters; the options are:\n");
  Config::printOptions();
}

static bool is(const char *a, const char *x, const char *y = nullptr) {
  return std::strcmp(a, x) == 0 || (y && std::strcmp(a, y) == 0);
}

template <class T>
static bool write_lines(const char *path, const std::vector<T> &v) {
  FILE *f = std::fopen(path, "w");
  if (!f) {
    std::printf("cannot write %s\n", path);
    return false;
  }
  for (const T &x : v) {
    if constexpr (std::is_floating_point<T>::value)
      std::fprintf(f, "%.9g\n", (double)x);
    else
      std::fprintf(f, "%d\n", (int)x);
  }
  std::fclose(f);
  return true;
}

static int run(int argc, char **argv) {
  const char *graph_file = nullptr, *out_file = nullptr, *sources_file = nullptr, *preds_file = nullptr;
  Config cfg;
  for (int i = 1; i < argc; i++) {
    const char *a = argv[i];
    if (is(a, "-help")) {
      usage(argv[0]);
      return 0;
    }
    if (i + 1 >= argc) {
      std::printf("option %s needs a value\n", a);
      return 1;
    }
    if (is(a, "-g", "-graph"))
      graph_file = argv[++i];
    else if (is(a, "-o", "-output"))
      out_file = argv[++i];
    else if (is(a, "-s", "-sources"))
      sources_file = argv[++i];
    else if (is(a, "-p", "-parameters"))
      cfg.parseParameterString(argv[++i]);
    else if (is(a, "-c"))
      cfg.parseFile(argv[++i]);
    else if (kMarkPredecessors && is(a, "-preds"))
      preds_file = argv[++i];
    else {
      std::printf("Wrong command: %s\n", a);
      return 1;
    }
  }
  if (!graph_file) {
    usage(argv[0]);
    return 1;
  }
  const int directed = cfg.getParameter<int>("directed");
  const int origin = cfg.getParameter<int>("origin");
  const int iter_num = cfg.getParameter<int>("iter_num");
  const int threshold = cfg.getParameter<int>("threshold");
  const int device = cfg.getParameter<int>("device");
  const bool verbose = cfg.getParameter<int>("verbose") != 0;

  CsrGraph<VertexId, Value, SizeT> csr_graph(false);
  if (builder::BuildMarketGraph<kWithValue>(const_cast<char *>(graph_file), csr_graph, false) != 0) return 1;
  const SizeT n = csr_graph.nodes;
  std::printf("%s: %d vertices, %d edges, %s\n", kProgram, (int)n, (int)csr_graph.edges,
              directed ? "directed" : "undirected");

  // Starting vertices (the programs that start from every vertex ignore them).
  std::vector<int> srcs;
  if (sources_file) {
    FILE *f = std::fopen(sources_file, "r");
    if (!f) {
      std::printf("Source file open error: %s\n", sources_file);
      return 1;
    }
    int v;
    while (std::fscanf(f, "%d", &v) == 1) srcs.push_back(origin == 1 ? v - 1 : v);
    std::fclose(f);
  } else {
    const int v = cfg.getParameter<int>("src");
    srcs.push_back(origin == 1 ? v - 1 : v);
  }
  if (Program::srcVertex() == ALL) srcs.assign(1, 0);
  for (int s : srcs) {
    if (s < 0 || s >= n) {
      std::printf("starting vertex %d out of range [0, %d)\n", s, (int)n);
      return 1;
    }
  }

  typedef GASengine::CsrProblem<Program, VertexId, SizeT, Value, kMarkPredecessors, kWithValue> Problem;
  Problem csr_problem(cfg);
  hipError_t e = csr_problem.FromHostProblem(false, csr_graph.nodes, csr_graph.edges, csr_graph.column_indices,
                                             csr_graph.row_offsets, csr_graph.edge_values, csr_graph.row_indices,
                                             csr_graph.column_offsets, 1, directed, device, 0);
  if (e != hipSuccess) {
    std::printf("FromHostProblem failed: %s\n", hipGetErrorString(e));
    return 1;
  }
  GASengine::EnactorVertexCentric<Problem, Program, true> vertex_centric(cfg, verbose);
  for (size_t i = 0; i < srcs.size(); i++) {
    int src[1] = {srcs[i]};
    if (Program::srcVertex() != ALL)
      std::printf("num_srcs=%d, src=%d, iter_num=%d\n", (int)srcs.size(), src[0], iter_num);
    e = vertex_centric.EnactIterativeSearch(csr_problem, csr_graph.row_offsets, directed, 1, src, iter_num,
                                            threshold);
    if (e != hipSuccess) {
      std::printf("EnactIterativeSearch failed: %s\n", hipGetErrorString(e));
      return 1;
    }
    std::printf("iterations %d, edges %lld\n", vertex_centric.iterations_run, (long long)vertex_centric.total_edges);
  }

  std::vector<Value> h_values((size_t)n);
  e = csr_problem.ExtractResults(h_values.data());
  if (e != hipSuccess) {
    std::printf("ExtractResults failed: %s\n", hipGetErrorString(e));
    return 1;
  }
  if (out_file && !write_lines(out_file, h_values)) return 1;
  if (kMarkPredecessors && preds_file) {
    std::vector<int> h_preds((size_t)n);
    e = csr_problem.ExtractPreds(h_preds.data());
    if (e != hipSuccess) {
      std::printf("ExtractPreds failed: %s\n", hipGetErrorString(e));
      return 1;
    }
    if (!write_lines(preds_file, h_preds)) return 1;
  }
  Program::Release(csr_problem.vertex_list, csr_problem.edge_list);
  return 0;
}

int main(int argc, char **argv) {
  try {
    return run(argc, argv);
  } catch (const std::invalid_argument &ex) {
    std::printf("%s\n", ex.what());
    return 1;
  }
}
