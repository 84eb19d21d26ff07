// serve.hip — serving loop: enumeration requests as they arrive, with
// continuous batching.  No reference counterpart.
//
// Requests are prepared on the submitting thread (a stand-alone graph is
// copied, an ego request is cut from the resident graph; both are brought to
// the -t 1 -o 1 -f working form) and queued.  A server thread owns several
// lanes (lane.hpp), each with its own stream and device buffers: whenever a
// lane is free and requests are waiting, it admits waiting requests in
// arrival order while the lane's slot count and memory budget allow, uploads
// them and launches one multi-graph run.  Lanes run concurrently, so requests
// that arrive while others are on the GPU start as soon as a lane frees up
// and their work overlaps on the device.  Completion is detected by polling
// the lane's event; results carry the submit and completion times.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "device_graph.hpp"
#include "graph.hpp"
#include "lane.hpp"

using namespace gmbe;

namespace {

constexpr int kLanes = 3;
constexpr uint32_t kMaxSide = 1u << 16;  // per-request side size the slabs are sized for

struct Request {
  uint64_t ticket = 0;
  std::shared_ptr<LanePack> one;  // the request in lane layout, built at submit
  uint32_t nU = 0, nV = 0;
  uint64_t E = 0;
  uint64_t bytes = 0;       // predicted device bytes
  double t_submit = 0;
};

// Device bytes a request occupies in a lane: CSR both ways, hash keys, table
// entry, claim-order entries and per-graph counters.
uint64_t predict_bytes(const gmbe_graph& g) {
  return 8ull * g.E + 16ull * ((uint64_t)g.nU + g.nV + 2) + sizeof(mbe::DevGraph) + 12 +
         8ull * mbe::kPerGraph;
}

// List memory a request surely needs in a wave's arena: the root list of a
// task holds one entry of 16 + 8 W bytes (W = bitset words of the root's L,
// at most 16) per 2-hop candidate, and a root v has at least deg(u) - 1 of
// them for each neighbour u.  The largest such lower bound over the roots;
// a request above the arena would overflow it.  O(E).
uint64_t predict_stack_bytes(const gmbe_graph& g) {
  uint64_t best = 0;
  for (uint32_t v = 0; v < g.nV; ++v) {
    uint64_t cand = 0;
    for (int64_t e = g.v_off[v]; e < g.v_off[v + 1]; ++e) {
      const uint32_t u = g.v_adj[e];
      cand = std::max<uint64_t>(cand, (uint64_t)(g.u_off[u + 1] - g.u_off[u]) - 1);
    }
    const uint64_t deg = (uint64_t)(g.v_off[v + 1] - g.v_off[v]);
    const uint64_t words = std::min<uint64_t>((deg + 63) / 64, (uint64_t)mbe::kMaxW);
    best = std::max<uint64_t>(best, cand * (16 + 8 * words));
  }
  return best;
}

}  // namespace

struct gmbe_server {
  gmbe_graph* resident = nullptr;  // file orientation (host copy for ego cuts)
  int slots = 0;
  uint64_t lane_bytes = 0;  // memory budget per lane
  int bh = 20, bs = 1500;
  Engine eng;
  Lane lanes[kLanes];
  std::vector<Request> in_lane[kLanes];
  bool busy[kLanes] = {false, false, false};
  LanePack pack;
  std::mutex mu;
  std::condition_variable cv, cv_done;
  std::deque<Request> pending;
  std::deque<gmbe_serve_done> done;
  uint64_t next_ticket = 1;
  int64_t request_ms = 0;  // deadline of each launch (0 = none)
  std::atomic<int64_t> stack_limit_kb{0};  // admission: predicted list memory per wave
  // statistics (written by the server thread under mu)
  uint64_t launches = 0, launched = 0;
  double host_s = 0, gpu_s = 0;
  uint64_t outstanding = 0;
  std::atomic<bool> stop{false};
  std::thread th;
  std::atomic<int> fatal{0};  // set by the server thread, read by submitters
};

namespace {

void finish_lane(gmbe_server* s, int l) {
  Lane& L = s->lanes[l];
  const double t = now_s();
  float ms = 0;
  if (hipEventElapsedTime(&ms, L.ev0, L.ev1) != hipSuccess) ms = 0;
  std::vector<gmbe_serve_done> out;
  for (size_t i = 0; i < s->in_lane[l].size(); ++i) {
    Request& r = s->in_lane[l][i];
    gmbe_result res;
    L.result((uint32_t)i, &res);
    gmbe_serve_done d{};
    d.ticket = r.ticket;
    d.status = res.overflow ? GMBE_ERR_CAPACITY : L.complete((uint32_t)i) ? GMBE_OK : GMBE_ERR_DEADLINE;
    d.maximal = res.maximal;
    d.hash_sum = res.hash_sum;
    d.hash_xor = res.hash_xor;
    d.t_submit_s = r.t_submit;
    d.t_done_s = t;
    out.push_back(d);
    r.one.reset();
  }
  s->in_lane[l].clear();
  s->busy[l] = false;
  std::lock_guard<std::mutex> lk(s->mu);
  s->gpu_s += ms * 1e-3;
  for (auto& d : out) s->done.push_back(d);
  s->outstanding -= out.size();
  s->cv_done.notify_all();
}

int launch_lane(gmbe_server* s, int l, std::vector<Request>& batch) {
  Lane& L = s->lanes[l];
  {
    std::lock_guard<std::mutex> lk(s->mu);
    L.knobs.budget_ms = s->request_ms;
  }
  const double t0 = now_s();
  std::vector<const LanePack*> parts;
  for (Request& r : batch) parts.push_back(r.one.get());
  // with a deadline, the cheap requests go first so that most of them finish
  int rc = L.upload_many(parts, L.knobs.budget_ms > 0);
  if (rc) return rc;
  if (hipEventRecord(L.ev0, L.st) != hipSuccess) return set_error(GMBE_ERR_HIP, "event record");
  {
    std::lock_guard<std::mutex> lk(s->mu);  // the statistics are read under the lock
    s->host_s += now_s() - t0;
    s->launches++;
    s->launched += batch.size();
  }
  rc = L.record_run();
  if (rc) return rc;
  if (hipEventRecord(L.ev1, L.st) != hipSuccess) return set_error(GMBE_ERR_HIP, "event record");
  // the pack's host copy is staged in the lane's pinned buffer: it can be reused now
  s->in_lane[l] = std::move(batch);
  s->busy[l] = true;
  return GMBE_OK;
}

void server_loop(gmbe_server* s) {
  (void)hipSetDevice(s->eng.device);
  while (!s->stop.load()) {
    bool progressed = false;
    for (int l = 0; l < kLanes; ++l) {
      if (s->busy[l] && hipEventQuery(s->lanes[l].ev1) == hipSuccess) {
        finish_lane(s, l);
        progressed = true;
      }
    }
    for (int l = 0; l < kLanes; ++l) {
      if (s->busy[l]) continue;
      std::vector<Request> batch;
      {
        std::lock_guard<std::mutex> lk(s->mu);
        const LaneCaps& c = s->lanes[l].caps;
        uint64_t voff = 0, uoff = 0, edges = 0, bytes = 0;
        while (!s->pending.empty() && (int)batch.size() < s->slots) {
          const Request& r = s->pending.front();
          if (bytes + r.bytes > s->lane_bytes || voff + r.nV + 1 > c.max_voff ||
              uoff + r.nU + 1 > c.max_uoff || edges + r.E > c.max_edges)
            break;  // admission: wait for memory to free up
          voff += r.nV + 1;
          uoff += r.nU + 1;
          edges += r.E;
          bytes += r.bytes;
          batch.push_back(s->pending.front());
          s->pending.pop_front();
        }
      }
      if (batch.empty()) break;
      if (launch_lane(s, l, batch) != GMBE_OK) {
        s->fatal.store(GMBE_ERR_HIP);
        std::lock_guard<std::mutex> lk(s->mu);
        for (Request& r : batch) {
          gmbe_serve_done d{};
          d.ticket = r.ticket;
          d.status = GMBE_ERR_HIP;
          d.t_submit_s = r.t_submit;
          d.t_done_s = now_s();
          s->done.push_back(d);
          --s->outstanding;
        }
        s->cv_done.notify_all();
      }
      progressed = true;
    }
    if (!progressed) {
      bool any_busy = false;
      for (int l = 0; l < kLanes; ++l) any_busy = any_busy || s->busy[l];
      std::unique_lock<std::mutex> lk(s->mu);
      if (any_busy) {
        // a lane is on the GPU: poll its event again shortly
        s->cv.wait_for(lk, std::chrono::microseconds(50));
      } else {
        s->cv.wait_for(lk, std::chrono::milliseconds(5),
                       [&] { return s->stop.load() || !s->pending.empty(); });
      }
    }
  }
  for (int l = 0; l < kLanes; ++l)
    if (s->busy[l]) {
      (void)hipStreamSynchronize(s->lanes[l].st);
      finish_lane(s, l);
    }
}

int submit(gmbe_server* s, gmbe_graph* w, uint64_t* ticket) {
  Request r;
  r.bytes = predict_bytes(*w);
  // admission by predicted list memory: a request whose root lists cannot fit
  // a wave's arena would overflow it; refuse it before it is queued
  const uint64_t stack = predict_stack_bytes(*w);
  const uint64_t limit = (uint64_t)s->stack_limit_kb.load() * 1024ull;
  if (stack > limit) {
    gmbe_graph_destroy(w);
    return set_error(GMBE_ERR_REJECTED,
                     "request needs %llu bytes of list memory per wave (predicted); the limit is "
                     "%llu", (unsigned long long)stack, (unsigned long long)limit);
  }
  r.t_submit = now_s();
  r.nU = w->nU;
  r.nV = w->nV;
  r.E = w->E;
  const LaneCaps& c = s->lanes[0].caps;
  if (r.bytes > s->lane_bytes || w->nV + 1 > c.max_voff || w->nU + 1 > c.max_uoff ||
      w->E > c.max_edges || w->nV > c.max_nV || w->nU >= kMaxSide) {
    gmbe_graph_destroy(w);
    return set_error(GMBE_ERR_REJECTED, "request of %llu predicted bytes does not fit the budget",
                     (unsigned long long)r.bytes);
  }
  uint32_t dmax = 0;
  for (uint32_t v = 0; v < w->nV; ++v)
    dmax = std::max<uint32_t>(dmax, (uint32_t)(w->v_off[v + 1] - w->v_off[v]));
  if (dmax > c.max_dmaxV) {
    gmbe_graph_destroy(w);
    return set_error(GMBE_ERR_REJECTED, "request degree %u exceeds the server's %u", dmax,
                     c.max_dmaxV);
  }
  // the lane layout (CSR copies, hash keys) is built here, on the caller's thread
  r.one = std::make_shared<LanePack>();
  r.one->add(*w, s->bh, s->bs);
  gmbe_graph_destroy(w);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    r.ticket = s->next_ticket++;
    ++s->outstanding;
    s->pending.push_back(r);
  }
  s->cv.notify_all();
  if (ticket) *ticket = r.ticket;
  return GMBE_OK;
}

}  // namespace

extern "C" {

static int gmbe_serve_create_impl(const gmbe_graph* resident, int slots, uint64_t mem_budget,
                                  int bound_height, int bound_size, gmbe_server** out) {
  if (!out) return set_error(GMBE_ERR_ARG, "null argument");
  *out = nullptr;
  if (slots < 1 || slots > 65536) return set_error(GMBE_ERR_ARG, "slots must be in [1, 65536]");
  if (bound_height < 0 || bound_size < 0) return set_error(GMBE_ERR_ARG, "bounds must be non-negative");
  if (resident && resident->swapped)
    return set_error(GMBE_ERR_ARG, "the resident graph must be in file orientation");
  gmbe_server* s = new gmbe_server();
  s->slots = slots;
  s->bh = bound_height;
  s->bs = bound_size;
  s->lane_bytes = mem_budget / kLanes;
  int rc = s->eng.ensure_device(0);
  if (rc == GMBE_OK && resident) {
    gmbe_option keep{0, 0, 1};
    rc = gmbe_graph_derive(resident, &keep, &s->resident);
  }
  // lane capacities from the per-lane budget (at least one small request)
  LaneCaps caps;
  caps.max_graphs = (uint32_t)slots;
  const uint64_t lb = std::max<uint64_t>(s->lane_bytes, 1);
  caps.max_edges = std::max<uint64_t>(lb / 12, 1);
  caps.max_voff = std::max<uint64_t>(lb / 16, 2) + slots;
  caps.max_uoff = caps.max_voff;
  caps.max_nV = kMaxSide;
  caps.max_dmaxV = kMaxSide;
  LaneKnobs kn;
  kn.arena_kb = 4096;
  kn.ctx_kb = 512;
  kn.queue_log2 = 18;
  kn.deferred_cap = 1ull << 16;
  s->stack_limit_kb = kn.arena_kb;
  for (int l = 0; l < kLanes && rc == GMBE_OK; ++l) {
    rc = s->lanes[l].alloc(&s->eng, caps, kn);
    if (rc == GMBE_OK) {
      // each lane gets half the resident grid, so that two launches share the GPU
      s->lanes[l].knobs.blocks = std::max(1, resident_blocks(s->lanes[l].kc, s->eng.cu_count) / 2);
      rc = s->lanes[l].alloc_run();
    }
  }
  if (rc) {
    gmbe_serve_destroy(s);
    return rc;
  }
  try {
    s->th = std::thread(server_loop, s);
  } catch (...) {
    gmbe_serve_destroy(s);  // no thread to join: releases the lanes only
    throw;                  // reported as an error code by the entry point's guard
  }
  *out = s;
  return GMBE_OK;
}

int gmbe_serve_create(const gmbe_graph* resident, int slots, uint64_t mem_budget, int bound_height,
                      int bound_size, gmbe_server** out) {
  return gmbe::guarded("gmbe_serve_create", [&] {
    return gmbe_serve_create_impl(resident, slots, mem_budget, bound_height, bound_size, out);
  });
}

static int gmbe_serve_set_impl(gmbe_server* s, const char* key, int64_t value) {
  if (!s || !key) return set_error(GMBE_ERR_ARG, "null argument");
  const std::string k(key);
  if (k == "stack_limit_kb") {
    // at most the lanes' arena: a larger limit would admit requests that overflow
    const int64_t arena = s->lanes[0].knobs.arena_kb;
    if (value < 1 || value > arena)
      return set_error(GMBE_ERR_ARG, "stack_limit_kb=%lld out of range [1, %lld]", (long long)value,
                       (long long)arena);
    s->stack_limit_kb = value;
    return GMBE_OK;
  }
  if (k == "request_ms") {
    if (value < 0 || value > (1ll << 32))
      return set_error(GMBE_ERR_ARG, "request_ms=%lld out of range", (long long)value);
    // read by the server thread when it records the next launch of each lane
    std::lock_guard<std::mutex> lk(s->mu);
    s->request_ms = value;
    return GMBE_OK;
  }
  return set_error(GMBE_ERR_ARG, "unknown knob '%s'", key);
}

int gmbe_serve_set(gmbe_server* s, const char* key, int64_t value) {
  return gmbe::guarded("gmbe_serve_set", [&] { return gmbe_serve_set_impl(s, key, value); });
}

static int gmbe_serve_stats_impl(gmbe_server* s, uint64_t* out, int n) {
  if (!s || (n > 0 && !out)) return set_error(GMBE_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  const uint64_t v[5] = {s->launches, s->launched, (uint64_t)(s->host_s * 1e6),
                         (uint64_t)(s->gpu_s * 1e6), (uint64_t)s->pending.size()};
  const int k = n < 5 ? n : 5;
  for (int i = 0; i < k; ++i) out[i] = v[i];
  return k;
}

int gmbe_serve_stats(gmbe_server* s, uint64_t* out, int n) {
  return gmbe::guarded("gmbe_serve_stats", [&] { return gmbe_serve_stats_impl(s, out, n); });
}

static int gmbe_serve_submit_graph_impl(gmbe_server* s, const gmbe_graph* g, uint64_t* ticket) {
  if (!s || !g) return set_error(GMBE_ERR_ARG, "null argument");
  if (s->fatal.load()) return set_error(s->fatal.load(), "the server has failed");
  gmbe_option run{1, 1, 1};
  gmbe_graph* w = nullptr;
  int rc = gmbe_graph_derive(g, &run, &w);
  if (rc) return rc;
  return submit(s, w, ticket);
}

int gmbe_serve_submit_graph(gmbe_server* s, const gmbe_graph* g, uint64_t* ticket) {
  return gmbe::guarded("gmbe_serve_submit_graph", [&] {
    return gmbe_serve_submit_graph_impl(s, g, ticket);
  });
}

static int gmbe_serve_submit_ego_impl(gmbe_server* s, uint32_t u, uint64_t max_edges,
                                      uint64_t* ticket) {
  if (!s) return set_error(GMBE_ERR_ARG, "null argument");
  if (!s->resident) return set_error(GMBE_ERR_STATE, "the server has no resident graph");
  if (s->fatal.load()) return set_error(s->fatal.load(), "the server has failed");
  gmbe_graph* ego = nullptr;
  int rc = gmbe_ego_subgraph(s->resident, u, max_edges, &ego);
  if (rc) return rc;
  gmbe_option run{1, 1, 1};
  gmbe_graph* w = nullptr;
  rc = gmbe_graph_derive(ego, &run, &w);
  gmbe_graph_destroy(ego);
  if (rc) return rc;
  return submit(s, w, ticket);
}

int gmbe_serve_submit_ego(gmbe_server* s, uint32_t u, uint64_t max_edges, uint64_t* ticket) {
  return gmbe::guarded("gmbe_serve_submit_ego", [&] {
    return gmbe_serve_submit_ego_impl(s, u, max_edges, ticket);
  });
}

static int gmbe_serve_poll_impl(gmbe_server* s, gmbe_serve_done* out, int max, int wait,
                                int* n_out) {
  if (!s || !n_out || (max > 0 && !out)) return set_error(GMBE_ERR_ARG, "null argument");
  std::unique_lock<std::mutex> lk(s->mu);
  if (wait && s->done.empty() && s->outstanding > 0)
    s->cv_done.wait(lk, [&] { return !s->done.empty() || s->outstanding == 0; });
  int n = 0;
  while (n < max && !s->done.empty()) {
    out[n++] = s->done.front();
    s->done.pop_front();
  }
  *n_out = n;
  return GMBE_OK;
}

int gmbe_serve_poll(gmbe_server* s, gmbe_serve_done* out, int max, int wait, int* n_out) {
  return gmbe::guarded("gmbe_serve_poll", [&] {
    return gmbe_serve_poll_impl(s, out, max, wait, n_out);
  });
}

void gmbe_serve_destroy(gmbe_server* s) {
  if (!s) return;
  s->stop.store(true);
  s->cv.notify_all();
  if (s->th.joinable()) s->th.join();
  for (int l = 0; l < kLanes; ++l) s->lanes[l].release();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->pending.clear();
  }
  if (s->resident) gmbe_graph_destroy(s->resident);
  delete s;
}

}  // extern "C"
