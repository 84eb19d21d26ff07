// gmbe.hpp — drop-in C++ classes with the reference's names, built on the C
// ABI of include/gmbe.h.  A program written against the reference's
// Option / CSRBiGraph / BicliqueFinder / IterFinderGpu6 (src/BiGraph.h:13-38,
// src/BicliqueFinder.h:18-30, src/IterFinderGpu.h:18-90) compiles against this
// header with the include changed.
//
// Differences, each deliberate:
//   * maximal_nodes_ is 64-bit (the reference's int wraps at 2^31, Q3);
//   * destructors are virtual (the reference deletes through a non-virtual
//     base destructor, src/main.cu:22, Q8);
//   * an error prints "GPUassert: <message>" and exits with the error code,
//     as gpuAssert does (src/Utility.h:30-39); the C ABI itself never exits;
//   * a missing file is an error instead of an empty graph (Q10).
#pragma once

#include <cstdint>
#include <cstdio>
#include <vector>
#include <cstdlib>
#include <cstring>

#include "../../include/gmbe.h"

inline void gmbe_assert(int rc, const char* what) {
  if (rc != GMBE_OK) {
    fprintf(stderr, "GPUassert: %s: %s\n", what, gmbe_last_error());
    exit(rc < 0 ? -rc : rc);
  }
}

struct Option {
  int uvtrans;
  int order;
  int fast_mode;
  Option() : uvtrans(0), order(1), fast_mode(0) {}
};

struct CSRBiGraph {
  CSRBiGraph() {}
  ~CSRBiGraph() { Reset(); }
  CSRBiGraph(const CSRBiGraph&) = delete;
  CSRBiGraph& operator=(const CSRBiGraph&) = delete;

  // CSRBiGraph::Read (src/BiGraph.cu:13-278) and its stats line
  // (src/BiGraph.cu:212-215, 247-248).
  void Read(const char* filename, Option opt) {
    Reset();
    gmbe_option o{opt.uvtrans, opt.order, opt.fast_mode};
    gmbe_assert(gmbe_graph_read_adj(filename, &o, &handle), "CSRBiGraph::Read");
    int64_t s[8];
    gmbe_graph_stats(handle, s);
    U_size = (int)s[0];
    V_size = (int)s[1];
    edge_size = (int)s[2];
    U_degree = (int)s[3];
    V_degree = (int)s[4];
    U_2degree = (int)s[5];
    V_2degree = (int)s[6];
    gmbe_graph_csr(handle, &row_offset, &column_indices, &rev_row_offset, &rev_column_indices,
                   nullptr, nullptr);
    printf("Graph:%s ", filename);
    if (opt.fast_mode) {
      printf("|U|:%8d |V|:%8d |E|:%8d\n", U_size, V_size, edge_size);
    } else {
      printf("|U|:%8d |V|:%8d |E|:%8d ", U_size, V_size, edge_size);
      printf("d(U):%8d d2(U):%8d ", U_degree, U_2degree);
      printf("d(V):%8d d2(V):%8d\n", V_degree, V_2degree);
    }
  }
  // The device copy is made by each finder (CopyToGpu, src/BiGraph.cu:413-431).
  void Reset() {
    if (handle) gmbe_graph_destroy(handle);
    handle = nullptr;
    U_size = V_size = edge_size = 0;
    row_offset = rev_row_offset = nullptr;
    column_indices = rev_column_indices = nullptr;
  }

  int U_size = 0, V_size = 0, edge_size = 0, U_degree = 0, V_degree = 0, U_2degree = 0,
      V_2degree = 0;
  const int64_t* row_offset = nullptr;        // V_size + 1
  const uint32_t* column_indices = nullptr;   // edge_size
  const int64_t* rev_row_offset = nullptr;    // U_size + 1
  const uint32_t* rev_column_indices = nullptr;
  gmbe_graph* handle = nullptr;
};

class BicliqueFinder {
 public:
  BicliqueFinder() = delete;
  explicit BicliqueFinder(CSRBiGraph* graph_in)
      : graph_(graph_in), maximal_nodes_(0), exe_time_(0), start_time_(0) {}
  virtual ~BicliqueFinder() {}
  virtual void Execute() = 0;
  // src/BicliqueFinder.cu:9-16
  void PrintResult(char* fn = nullptr) {
    FILE* fp = (fn == nullptr || strlen(fn) == 0) ? stdout : fopen(fn, "a+");
    if (!fp) fp = stdout;
    if (fp != stdout) fseek(fp, 0, SEEK_END);
    fprintf(fp, "Total processing time: %lf seconds\n", exe_time_);
    fprintf(fp, "maximal nodes:%llu\n", (unsigned long long)maximal_nodes_);
    fprintf(fp, "\n");
    if (fp != stdout) fclose(fp);
  }
  uint64_t MaximalNodes() const { return maximal_nodes_; }
  double ExeTime() const { return exe_time_; }

 protected:
  CSRBiGraph* graph_;
  uint64_t maximal_nodes_;
  double exe_time_, start_time_;
};

// The hot path: IterFinderGpu6 (src/IterFinderGpu.h:77-90).  sel picks the
// scheduler; the reference's IterFinderGpu (sel 0) and IterFinderGpu2 (sel 1)
// map to the same kernel without task splitting.
// -p of the reference's main (src/main.cu:121-123): print each block's exit
// time after a run ("SM exit", src/IterFinderGpu.cu:1038-1041).
inline bool& print_exit_times() {
  static bool on = false;
  return on;
}

class IterFinderGpu6 : public BicliqueFinder {
 public:
  IterFinderGpu6() = delete;
  IterFinderGpu6(CSRBiGraph* graph_in, int bound_height = 20, int bound_size = 1500, int sel = 2)
      : BicliqueFinder(graph_in), bound_height_(bound_height), bound_size_(bound_size) {
    gmbe_assert(gmbe_finder_create(graph_in->handle, sel, bound_height, bound_size, 1, &finder_),
                "IterFinderGpu6");
  }
  ~IterFinderGpu6() override { gmbe_finder_destroy(finder_); }
  // src/IterFinderGpu.cu:1095-1119, same stdout lines
  void Execute() override {
    printf("large worklist size: %d local tiny size: %d\n", 1 << 22, 0);
    if (print_exit_times()) gmbe_finder_set(finder_, "exit_stamps", 1);
    gmbe_result r;
    gmbe_assert(gmbe_finder_execute(finder_, &r), "IterFinderGpu6::Execute");
    if (print_exit_times()) {
      size_t n = 0;
      if (gmbe_finder_exit_times(finder_, nullptr, 0, &n) == 0 && n > 0) {
        std::vector<double> t(n);
        gmbe_finder_exit_times(finder_, t.data(), n, &n);
        for (double x : t) printf("SM exit: %lf\n", x);
      }
    }
    printf("%d\n", 0);
    printf("non_maximal: %lld\n", (long long)r.non_maximal);
    maximal_nodes_ = r.maximal;
    exe_time_ = r.exe_time_s;
    last_ = r;
  }
  const gmbe_result& Result() const { return last_; }
  gmbe_finder* Handle() { return finder_; }

 protected:
  gmbe_finder* finder_ = nullptr;
  int bound_height_, bound_size_;
  gmbe_result last_{};
};

class IterFinderGpu : public IterFinderGpu6 {
 public:
  explicit IterFinderGpu(CSRBiGraph* g) : IterFinderGpu6(g, 20, 1500, 0) {}
};
class IterFinderGpu2 : public IterFinderGpu6 {
 public:
  explicit IterFinderGpu2(CSRBiGraph* g) : IterFinderGpu6(g, 20, 1500, 1) {}
};
