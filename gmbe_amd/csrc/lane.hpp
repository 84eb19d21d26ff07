// lane.hpp — device buffers for one multi-graph launch at a time.
//
// A lane holds a set of graphs in one concatenated device CSR with a graph
// table (one mbe::DevGraph per graph, pointing into the concatenation), the
// order in which their top-level vertices are claimed, per-graph result
// counters, and the run buffers of the persistent enumeration kernel
// (per-wave slabs, task queue, control block, deferred tasks).  The batched
// form replays one lane through a captured HIP graph; the serving loop keeps
// several lanes in flight on their own streams so that the launches overlap.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "device_graph.hpp"
#include "graph.hpp"
#include "mbe_kernel.hpp"

namespace gmbe {

// Capacities fixed at allocation.
struct LaneCaps {
  uint32_t max_graphs = 0;
  uint64_t max_voff = 0;   // sum over graphs of (nV + 1)
  uint64_t max_uoff = 0;   // sum over graphs of (nU + 1)
  uint64_t max_edges = 0;  // sum over graphs of E
  uint32_t max_nV = 0, max_dmaxV = 0;
};

// Kernel knobs of a lane (shared with the finder's meaning).
struct LaneKnobs {
  int lds_kb = 7, waves = 4, blocks = 0;
  int64_t arena_kb = 2048, ctx_kb = 1024, ctx_slots = 4, queue_log2 = 20;
  int prune = 1, max_prefix = 4, group = 1, share = 1, split = 1;
  int64_t share_work = 4096, share_queue = -1, merge_min = 256, merge_branch = 16,
          late_share = 32;
  int dom_k = 6, dom_skip = 1;  // list reduction (reduce_list)
  int recompute = 0;            // split large tasks without a free context into children
  int count_bytes = 0;          // count algorithmic bytes (alg_bytes)
  uint64_t deferred_cap = 1ull << 18;
  int64_t budget_ms = 0;  // per-wave deadline of a run (0 = none; partial results)
};

// Host copy of a set of graphs in lane layout.
struct LanePack {
  std::vector<int64_t> voff, uoff;
  std::vector<uint32_t> vadj, uadj;
  std::vector<uint64_t> kU, kV;
  std::vector<mbe::DevGraph> table;  // pointers are offsets until placed
  std::vector<double> cost;          // claim-order cost proxy per graph
  std::vector<uint32_t> order;       // claim order of the graphs
  std::vector<unsigned long long> root_base;
  // with size classes: order position of each class's first graph (plus the
  // end), and root_base holds one segment per class (n_c + 1 entries from 0)
  std::vector<uint32_t> seg_first;
  std::vector<int> seg_class;
  uint32_t max_nV = 0, max_dmaxV = 0;
  // add graph g with split thresholds (bh, bs); returns its index
  uint32_t add(const gmbe_graph& g, int bh, int bs);
  // append every graph of `o` (prepared elsewhere, e.g. on another thread)
  void append(const LanePack& o);
  // claim order by a cost proxy (sum of squared U degrees): heaviest first
  // for throughput, lightest first when requests race a deadline
  void finish(bool light_first = false);
  // claim order by size class (cls[i] of graph i, in [0, n_classes)), then
  // by cost, heaviest first; the kernel of each class runs over its own
  // segment of the order
  void finish_classes(const std::vector<int>& cls, int n_classes);
  void clear();
  uint64_t total_roots() const;
};

// Claim range of one size class in a lane (finish_classes).
struct LaneSeg {
  uint32_t g0 = 0, n = 0;  // first order position, graphs
  uint64_t rb_off = 0;     // its root_base segment
  uint64_t roots = 0;
  int cls = 0;
};

constexpr int kLaneClasses = 4;

struct Lane {
  Engine* eng = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  LaneCaps caps;
  LaneKnobs knobs;
  // device graph storage
  int64_t *voff = nullptr, *uoff = nullptr;
  uint32_t *vadj = nullptr, *uadj = nullptr;
  uint64_t *kU = nullptr, *kV = nullptr;
  mbe::DevGraph* gt = nullptr;
  unsigned long long* root_base = nullptr;
  uint32_t* order = nullptr;
  unsigned long long* per_graph = nullptr;
  // run buffers
  char* scratch = nullptr;
  uint64_t scratch_bytes = 0;
  mbe::Scratch layout{};
  int nwaves = 0, nblocks = 0, wblocks = 0;
  uint32_t* q = nullptr;
  unsigned long long* qseq = nullptr;
  uint64_t qcap = 0;
  unsigned long long* ctl = nullptr;
  unsigned long long* acc = nullptr;
  mbe::Task* deferred = nullptr;
  // pinned host staging: uploads, then results
  char* stage = nullptr;
  uint64_t stage_bytes = 0;
  unsigned long long* out_host = nullptr;  // per_graph (max_graphs * kPerGraph) + acc
  // current contents
  uint32_t n_graphs = 0;
  uint64_t n_roots = 0;
  std::vector<LaneSeg> segs;  // per-class claim ranges (empty: one range)
  KernelChoice kc{}, wk{};
  // per-class kernel variants (used when the contents have class segments):
  // waves per block and LDS KB per wave; -1 = the lane's knobs
  int cls_waves[kLaneClasses] = {-1, -1, -1, -1}, cls_lds[kLaneClasses] = {-1, -1, -1, -1};

  int alloc(Engine* e, const LaneCaps& c, const LaneKnobs& k);
  // (re)size the run buffers for the current knobs (after a knob change)
  int alloc_run();
  // kernel variant of class c (its own, else the knobs'), its block count
  KernelChoice class_kernel(int c, int* blocks) const;
  void release();
  // Copy a pack into the lane (async on st; the pack must stay alive until
  // the stream has passed the copy).
  int upload(const LanePack& p);
  // The same for the union of several packs (one graph each or more), copied
  // straight into the staging buffer; claim order by the packs' cost proxy.
  int upload_many(const std::vector<const LanePack*>& parts, bool light_first);
  // Record one full run on st: control reset, main pass, wide pass, results
  // to out_host.  No memset nodes, no host synchronisation: capturable.
  int record_run();
  // Kernel parameters of the main and wide passes.
  mbe::Params params(bool wide) const;
  // Per-graph result i (valid after the stream has passed record_run).
  void result(uint32_t i, gmbe_result* r) const;
  // Whether graph i finished (false only for a run cut by its deadline).
  bool complete(uint32_t i) const;
  // Totals of the last run.
  void totals(gmbe_result* r) const;
};

}  // namespace gmbe
