// batch.hip — batched form: many small graphs in one captured run (one
// persistent launch pair per graph size class), replayed from a HIP graph,
// with an autotuner that picks each size class's kernel variant (waves per
// block, LDS list arena per wave) and split thresholds.  No reference counterpart: the
// reference runs one graph per process (src/main.cu:126-147).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "device_graph.hpp"
#include "graph.hpp"
#include "lane.hpp"

using namespace gmbe;

namespace {

constexpr int kClasses = kLaneClasses;
// size classes by edge count: [0, 2k), [2k, 8k), [8k, 32k), [32k, inf)
int size_class(uint64_t E) { return E < 2000 ? 0 : E < 8000 ? 1 : E < 32000 ? 2 : 3; }

// kernel variants the autotuner tries: (waves per block, LDS KB per wave)
constexpr int kVariants[][2] = {{4, 7}, {4, 0}, {4, 3}, {4, 16}, {8, 7}, {2, 7}};
// split-threshold candidates per class: (bound_height, bound_size)
constexpr int kBounds[][2] = {{20, 1500}, {1 << 30, 1 << 30}, {10, 500}, {40, 6000}};

}  // namespace

struct gmbe_batch {
  std::vector<const gmbe_graph*> graphs;
  std::vector<int> cls;
  int cls_bh[kClasses], cls_bs[kClasses];
  Engine eng;
  LaneKnobs knobs;
  Lane lane;
  LanePack pack;
  bool packed = false;       // the lane holds the current pack and thresholds
  int64_t per_graph_launch = 0;  // 1: one launch per graph (the plain-kernel loop)
  int64_t verbose = 0;           // 1: the autotuner reports each trial on stderr
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool captured = false;
  double last_kernel_s = 0, last_wall_s = 0;
};

namespace {

void drop_capture(gmbe_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->captured = false;
}

// Load the lane with the graphs `subset` (all graphs when empty), in one
// claim segment per size class.
int repack(gmbe_batch* b, const std::vector<size_t>& subset = {}) {
  b->pack.clear();
  const size_t n = subset.empty() ? b->graphs.size() : subset.size();
  std::vector<int> cls(n);
  for (size_t j = 0; j < n; ++j) {
    const size_t i = subset.empty() ? j : subset[j];
    const int c = b->cls[i];
    b->pack.add(*b->graphs[i], b->cls_bh[c], b->cls_bs[c]);
    cls[j] = c;
  }
  b->pack.finish_classes(cls, kClasses);
  int rc = b->lane.upload(b->pack);
  if (rc) return rc;
  if (hipStreamSynchronize(b->lane.st) != hipSuccess)
    return set_error(GMBE_ERR_HIP, "batch upload failed");
  b->packed = subset.empty();
  drop_capture(b);
  return GMBE_OK;
}

int capture(gmbe_batch* b) {
  drop_capture(b);
  if (hipStreamBeginCapture(b->lane.st, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return set_error(GMBE_ERR_HIP, "stream capture could not start");
  int rc = b->lane.record_run();
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->lane.st, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return set_error(GMBE_ERR_HIP, "stream capture failed: %s", hipGetErrorString(e));
  b->graph = g;
  e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
  if (e != hipSuccess)
    return set_error(GMBE_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(e));
  b->captured = true;
  return GMBE_OK;
}

// One whole-batch run through the captured graph.
int run_captured(gmbe_batch* b) {
  if (!b->packed) {
    int rc = repack(b);
    if (rc) return rc;
  }
  if (!b->captured) {
    int rc = capture(b);
    if (rc) return rc;
  }
  Lane& L = b->lane;
  const double t0 = now_s();
  if (hipEventRecord(L.ev0, L.st) != hipSuccess || hipGraphLaunch(b->exec, L.st) != hipSuccess ||
      hipEventRecord(L.ev1, L.st) != hipSuccess || hipStreamSynchronize(L.st) != hipSuccess)
    return set_error(GMBE_ERR_HIP, "batch launch failed: %s", hipGetErrorString(hipGetLastError()));
  b->last_wall_s = now_s() - t0;
  float ms = 0;
  (void)hipEventElapsedTime(&ms, L.ev0, L.ev1);
  b->last_kernel_s = ms * 1e-3;
  return GMBE_OK;
}

// The plain-kernel loop: graph i alone, one launch and one host wait each.
int run_looped(gmbe_batch* b, std::vector<gmbe_result>& out) {
  drop_capture(b);
  b->packed = false;
  out.assign(b->graphs.size(), gmbe_result{});
  const double t0 = now_s();
  double kern = 0;
  LanePack one;
  for (size_t i = 0; i < b->graphs.size(); ++i) {
    one.clear();
    const int c = b->cls[i];
    one.add(*b->graphs[i], b->cls_bh[c], b->cls_bs[c]);
    one.finish_classes({c}, kClasses);  // the graph's class kernel
    int rc = b->lane.upload(one);
    if (rc) return rc;
    if (hipEventRecord(b->lane.ev0, b->lane.st) != hipSuccess)
      return set_error(GMBE_ERR_HIP, "event record failed");
    rc = b->lane.record_run();
    if (rc) return rc;
    if (hipEventRecord(b->lane.ev1, b->lane.st) != hipSuccess ||
        hipStreamSynchronize(b->lane.st) != hipSuccess)
      return set_error(GMBE_ERR_HIP, "looped launch failed");
    float ms = 0;
    (void)hipEventElapsedTime(&ms, b->lane.ev0, b->lane.ev1);
    kern += ms * 1e-3;
    b->lane.result(0, &out[i]);
    if (!b->lane.complete(0))
      return set_error(GMBE_ERR_STATE, "graph %zu has unfinished work after the run", i);
  }
  b->last_wall_s = now_s() - t0;
  b->last_kernel_s = kern;
  return GMBE_OK;
}

// The kernel variant of every class (c < 0) or of class c.
int set_variant(gmbe_batch* b, int waves, int lds_kb, int c = -1) {
  if (!pick_kernel(lds_kb, waves, false).fn)
    return set_error(GMBE_ERR_ARG, "no kernel with %d waves per block and lds_kb=%d", waves,
                     lds_kb);
  if (c < 0) {
    b->lane.knobs.waves = waves;
    b->lane.knobs.lds_kb = lds_kb;
    for (int k = 0; k < kClasses; ++k) b->lane.cls_waves[k] = b->lane.cls_lds[k] = -1;
  } else {
    b->lane.cls_waves[c] = waves;
    b->lane.cls_lds[c] = lds_kb;
  }
  drop_capture(b);
  return b->lane.alloc_run();
}
int class_waves(const gmbe_batch* b, int c) {
  return b->lane.cls_waves[c] > 0 ? b->lane.cls_waves[c] : b->lane.knobs.waves;
}
int class_lds(const gmbe_batch* b, int c) {
  return b->lane.cls_lds[c] >= 0 ? b->lane.cls_lds[c] : b->lane.knobs.lds_kb;
}

// Time one trial run (median of `reps`), recorded eagerly under a deadline:
// a run that reaches it counts as slower than any finished one.
int timed(gmbe_batch* b, const std::vector<size_t>& sample, int reps, double budget_s,
          double* sec) {
  int rc = repack(b, sample);
  if (rc) return rc;
  Lane& L = b->lane;
  L.knobs.budget_ms = (int64_t)(budget_s * 1e3) + 1;
  std::vector<double> t;
  for (int r = 0; r < reps && rc == GMBE_OK; ++r) {
    const double t0 = now_s();
    rc = L.record_run();
    if (rc == GMBE_OK && hipStreamSynchronize(L.st) != hipSuccess)
      rc = set_error(GMBE_ERR_HIP, "autotuner run failed");
    const double w = now_s() - t0;
    bool cut = w >= 0.95 * budget_s;
    for (uint32_t i = 0; i < L.n_graphs && !cut; ++i)
      cut = L.out_host[(uint64_t)i * mbe::kPerGraph + mbe::G_OVERFLOW] != 0;
    t.push_back(cut ? 1e30 : w);
  }
  L.knobs.budget_ms = 0;
  if (rc) return rc;
  *sec = *std::min_element(t.begin(), t.end());
  return GMBE_OK;
}

// Autotuner: for each size class, its kernel variant and then its split
// thresholds, each by the measured wall time of a sample of the class's
// graphs (best of two runs; slow candidates are cut short by the kernel's
// deadline).  The captured run is rebuilt afterwards.
int autotune(gmbe_batch* b) {
  // per class: up to kSample graphs spread over the class's members
  constexpr size_t kSample = 8;
  // trials are cut at twice the best time so far (at least 200 ms)
  auto budget = [&](double cur) { return std::max(0.2, 2.0 * std::min(cur, 1e3)); };
  for (int c = 0; c < kClasses; ++c) {
    std::vector<size_t> mem, sample;
    for (size_t i = 0; i < b->cls.size(); ++i)
      if (b->cls[i] == c) mem.push_back(i);
    if (mem.empty()) continue;
    const size_t step = std::max<size_t>(1, mem.size() / kSample);
    for (size_t j = 0; j < mem.size() && j / step < kSample; j += step) sample.push_back(mem[j]);
    double best = 1e300;
    int bw = class_waves(b, c), bl = class_lds(b, c);
    for (const auto& v : kVariants) {
      int rc = set_variant(b, v[0], v[1], c);
      if (rc) return rc;
      double s = 0;
      rc = timed(b, sample, 2, budget(best), &s);
      if (rc) return rc;
      if (b->verbose) {
        fprintf(stderr, "autotune: class %d, %zu graphs, waves_per_block=%d lds_kb=%d: %.4f s\n",
                c, sample.size(), v[0], v[1], s);
        fflush(stderr);
      }
      if (s < best) {
        best = s;
        bw = v[0];
        bl = v[1];
      }
    }
    int rc = set_variant(b, bw, bl, c);
    if (rc) return rc;
    int bh = b->cls_bh[c], bs = b->cls_bs[c];
    double cbest = 1e300;
    for (const auto& t : kBounds) {
      b->cls_bh[c] = t[0];
      b->cls_bs[c] = t[1];
      double s = 0;
      rc = timed(b, sample, 2, budget(std::min(cbest, best)), &s);
      if (rc) return rc;
      if (b->verbose) {
        fprintf(stderr, "autotune: class %d bound_height=%d bound_size=%d: %.4f s\n", c, t[0],
                t[1], s);
        fflush(stderr);
      }
      if (s < cbest) {
        cbest = s;
        bh = t[0];
        bs = t[1];
      }
    }
    b->cls_bh[c] = bh;
    b->cls_bs[c] = bs;
  }
  b->packed = false;  // the next execute loads the whole batch again
  return GMBE_OK;
}

}  // namespace

extern "C" {

static int gmbe_batch_create_impl(const gmbe_graph* const* graphs, int n, int bound_height,
                                  int bound_size, gmbe_batch** out) {
  if (!out || (n > 0 && !graphs)) return set_error(GMBE_ERR_ARG, "null argument");
  *out = nullptr;
  if (n < 1 || (uint32_t)n > mbe::kMaxGraphs)
    return set_error(GMBE_ERR_ARG, "batch size %d out of range [1, %u]", n, mbe::kMaxGraphs);
  if (bound_height < 0 || bound_size < 0) return set_error(GMBE_ERR_ARG, "bounds must be non-negative");
  gmbe_batch* b = new gmbe_batch();
  LaneCaps caps;
  caps.max_graphs = (uint32_t)n;
  for (int i = 0; i < n; ++i) {
    const gmbe_graph* g = graphs[i];
    if (!g) {
      delete b;
      return set_error(GMBE_ERR_ARG, "graph %d is null", i);
    }
    b->graphs.push_back(g);
    b->cls.push_back(size_class(g->E));
    caps.max_voff += g->nV + 1;
    caps.max_uoff += g->nU + 1;
    caps.max_edges += g->E;
    caps.max_nV = std::max(caps.max_nV, g->nV);
    for (uint32_t v = 0; v < g->nV; ++v)
      caps.max_dmaxV = std::max<uint32_t>(caps.max_dmaxV, (uint32_t)(g->v_off[v + 1] - g->v_off[v]));
  }
  for (int c = 0; c < kClasses; ++c) {
    b->cls_bh[c] = bound_height;
    b->cls_bs[c] = bound_size;
  }
  int rc = b->eng.ensure_device(0);
  if (rc == GMBE_OK) rc = b->lane.alloc(&b->eng, caps, b->knobs);
  if (rc) {
    gmbe_batch_destroy(b);
    return rc;
  }
  *out = b;
  return GMBE_OK;
}

int gmbe_batch_create(const gmbe_graph* const* graphs, int n, int bound_height, int bound_size,
                      gmbe_batch** out) {
  return gmbe::guarded("gmbe_batch_create", [&] {
    return gmbe_batch_create_impl(graphs, n, bound_height, bound_size, out);
  });
}

static int gmbe_batch_set_impl(gmbe_batch* b, const char* key, int64_t value) {
  if (!b || !key) return set_error(GMBE_ERR_ARG, "null argument");
  const std::string k(key);
  auto range = [&](int64_t lo, int64_t hi) -> int {
    if (value < lo || value > hi)
      return set_error(GMBE_ERR_ARG, "%s=%lld out of range [%lld, %lld]", key, (long long)value,
                       (long long)lo, (long long)hi);
    return GMBE_OK;
  };
  int rc;
  LaneKnobs& kn = b->lane.knobs;
  if (k.rfind("bound_height.", 0) == 0 || k.rfind("bound_size.", 0) == 0) {
    const bool h = k[6] == 'h';
    const int c = atoi(k.c_str() + (h ? 13 : 11));
    if (c < 0 || c >= kClasses) return set_error(GMBE_ERR_ARG, "size class %d out of range", c);
    if ((rc = range(0, 1 << 30))) return rc;
    (h ? b->cls_bh : b->cls_bs)[c] = (int)value;
    b->packed = false;
    return GMBE_OK;
  }
  if (k == "autotune") {
    if ((rc = range(0, 1))) return rc;
    return value ? autotune(b) : GMBE_OK;
  }
  if (k == "verbose") {
    if ((rc = range(0, 1))) return rc;
    b->verbose = value;
    return GMBE_OK;
  }
  if (k == "per_graph_launch") {
    if ((rc = range(0, 1))) return rc;
    b->per_graph_launch = value;
    return GMBE_OK;
  }
  if (k == "waves_per_block" || k.rfind("waves_per_block.", 0) == 0) {
    // every class, or class c with "waves_per_block.c"
    const int c = k.size() > 15 ? atoi(k.c_str() + 16) : -1;
    if (k.size() > 15 && (c < 0 || c >= kClasses))
      return set_error(GMBE_ERR_ARG, "size class %d out of range", c);
    if ((rc = range(2, 8))) return rc;
    // keep the LDS arena if that variant is built, else take the one that is
    int lds = c < 0 ? kn.lds_kb : class_lds(b, c);
    if (!pick_kernel(lds, (int)value, false).fn) lds = 7;
    return set_variant(b, (int)value, lds, c);
  }
  if (k == "lds_kb" || k.rfind("lds_kb.", 0) == 0) {
    const int c = k.size() > 6 ? atoi(k.c_str() + 7) : -1;
    if (k.size() > 6 && (c < 0 || c >= kClasses))
      return set_error(GMBE_ERR_ARG, "size class %d out of range", c);
    if ((rc = range(0, 24))) return rc;
    return set_variant(b, c < 0 ? kn.waves : class_waves(b, c), (int)value, c);
  }
  int64_t* slot = nullptr;
  int islot = -1;
  int* ip = nullptr;
  if (k == "prune") { if ((rc = range(0, 1))) return rc; ip = &kn.prune; }
  else if (k == "max_prefix") { if ((rc = range(2, mbe::kMaxPrefix))) return rc; ip = &kn.max_prefix; }
  else if (k == "group") { if ((rc = range(0, 1))) return rc; ip = &kn.group; }
  else if (k == "share") { if ((rc = range(0, 1))) return rc; ip = &kn.share; }
  else if (k == "split") { if ((rc = range(0, 1))) return rc; ip = &kn.split; }
  else if (k == "dom_k") { if ((rc = range(0, mbe::kMaxDom))) return rc; ip = &kn.dom_k; }
  else if (k == "dom_skip") { if ((rc = range(0, 1))) return rc; ip = &kn.dom_skip; }
  else if (k == "recompute") { if ((rc = range(0, 1))) return rc; ip = &kn.recompute; }
  else if (k == "count_bytes") { if ((rc = range(0, 1))) return rc; ip = &kn.count_bytes; }
  else if (k == "share_work") { if ((rc = range(1, 1ll << 40))) return rc; slot = &kn.share_work; }
  else if (k == "merge_min") { if ((rc = range(2, 1ll << 30))) return rc; slot = &kn.merge_min; }
  else if (k == "merge_branch") { if ((rc = range(0, 1ll << 30))) return rc; slot = &kn.merge_branch; }
  else if (k == "late_share") { if ((rc = range(0, 1ll << 30))) return rc; slot = &kn.late_share; }
  else if (k == "arena_kb") { if ((rc = range(64, 1 << 20))) return rc; slot = &kn.arena_kb; islot = 1; }
  else if (k == "ctx_kb") { if ((rc = range(16, 1 << 20))) return rc; slot = &kn.ctx_kb; islot = 1; }
  else if (k == "ctx_slots") { if ((rc = range(1, 32))) return rc; slot = &kn.ctx_slots; islot = 1; }
  else return set_error(GMBE_ERR_ARG, "unknown knob '%s'", key);
  if (ip) *ip = (int)value;
  if (slot) *slot = value;
  drop_capture(b);
  if (islot == 1) return b->lane.alloc_run();
  return GMBE_OK;
}

int gmbe_batch_set(gmbe_batch* b, const char* key, int64_t value) {
  return gmbe::guarded("gmbe_batch_set", [&] { return gmbe_batch_set_impl(b, key, value); });
}

static int gmbe_batch_get_impl(const gmbe_batch* b, const char* key, int64_t* value) {
  if (!b || !key || !value) return set_error(GMBE_ERR_ARG, "null argument");
  const std::string k(key);
  if (k == "waves_per_block") *value = b->lane.knobs.waves;
  else if (k == "lds_kb") *value = b->lane.knobs.lds_kb;
  else if (k.rfind("waves_per_block.", 0) == 0 || k.rfind("lds_kb.", 0) == 0) {
    const bool w = k[0] == 'w';
    const int c = atoi(k.c_str() + (w ? 16 : 7));
    if (c < 0 || c >= kClasses) return set_error(GMBE_ERR_ARG, "size class %d out of range", c);
    *value = w ? class_waves(b, c) : class_lds(b, c);
  }
  else if (k == "n_classes") *value = kClasses;
  else if (k == "kernel_time_us") *value = (int64_t)(b->last_kernel_s * 1e6);
  else if (k == "wall_time_us") *value = (int64_t)(b->last_wall_s * 1e6);
  else if (k.rfind("bound_height.", 0) == 0 || k.rfind("bound_size.", 0) == 0) {
    const bool h = k[6] == 'h';
    const int c = atoi(k.c_str() + (h ? 13 : 11));
    if (c < 0 || c >= kClasses) return set_error(GMBE_ERR_ARG, "size class %d out of range", c);
    *value = (h ? b->cls_bh : b->cls_bs)[c];
  } else if (k.rfind("class.", 0) == 0) {
    const long i = atol(k.c_str() + 6);
    if (i < 0 || (size_t)i >= b->cls.size()) return set_error(GMBE_ERR_ARG, "graph %ld out of range", i);
    *value = b->cls[i];
  } else {
    return set_error(GMBE_ERR_ARG, "unknown key '%s'", key);
  }
  return GMBE_OK;
}

int gmbe_batch_get(const gmbe_batch* b, const char* key, int64_t* value) {
  return gmbe::guarded("gmbe_batch_get", [&] { return gmbe_batch_get_impl(b, key, value); });
}

static int gmbe_batch_execute_impl(gmbe_batch* b, gmbe_result* per_graph, gmbe_result* total) {
  if (!b) return set_error(GMBE_ERR_ARG, "null argument");
  if (hipSetDevice(b->eng.device) != hipSuccess) return set_error(GMBE_ERR_HIP, "hipSetDevice");
  std::vector<gmbe_result> res;
  int rc;
  if (b->per_graph_launch) {
    rc = run_looped(b, res);
    if (rc) return rc;
  } else {
    rc = run_captured(b);
    if (rc) return rc;
    res.resize(b->graphs.size());
    for (size_t i = 0; i < b->graphs.size(); ++i) {
      b->lane.result((uint32_t)i, &res[i]);
      // every unit of every graph accounted for: a run without deadline ends complete
      if (!b->lane.complete((uint32_t)i))
        return set_error(GMBE_ERR_STATE, "graph %zu has unfinished work after the run", i);
    }
  }
  gmbe_result t{};
  for (const gmbe_result& r : res) {
    t.maximal += r.maximal;
    t.hash_sum += r.hash_sum;
    t.hash_xor ^= r.hash_xor;
    t.non_maximal += r.non_maximal;
    t.nodes += r.nodes;
    t.tasks += r.tasks;
    t.overflow += r.overflow;
    t.alg_bytes += r.alg_bytes;
  }
  t.exe_time_s = b->last_wall_s;
  t.kernel_time_s = b->last_kernel_s;
  if (per_graph) memcpy(per_graph, res.data(), res.size() * sizeof(gmbe_result));
  if (total) *total = t;
  if (t.overflow)
    return set_error(GMBE_ERR_CAPACITY, "%llu task(s) exceeded a capacity; the result is incomplete",
                     (unsigned long long)t.overflow);
  return GMBE_OK;
}

int gmbe_batch_execute(gmbe_batch* b, gmbe_result* per_graph, gmbe_result* total) {
  return gmbe::guarded("gmbe_batch_execute", [&] {
    return gmbe_batch_execute_impl(b, per_graph, total);
  });
}

void gmbe_batch_destroy(gmbe_batch* b) {
  if (!b) return;
  drop_capture(b);
  b->lane.release();
  delete b;
}

}  // extern "C"
