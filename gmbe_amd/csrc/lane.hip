// lane.hip — device buffers and launch sequence of one multi-graph run.
#include "lane.hpp"

#include <algorithm>
#include <cstring>

namespace gmbe {

#define LANE_TRY(expr)                                                                       \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return set_error(GMBE_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                       __FILE__, __LINE__);                                                  \
  } while (0)

using mbe::kCtlStride;
using mbe::kPerGraph;

// ---------------------------------------------------------------------------
// LanePack

uint32_t LanePack::add(const gmbe_graph& g, int bh, int bs) {
  const uint32_t id = (uint32_t)table.size();
  mbe::DevGraph d{};
  // offsets into the concatenated arrays, turned into pointers by Lane::upload
  d.v_off = (const int64_t*)(uintptr_t)voff.size();
  d.v_adj = (const uint32_t*)(uintptr_t)vadj.size();
  d.u_off = (const int64_t*)(uintptr_t)uoff.size();
  d.u_adj = (const uint32_t*)(uintptr_t)uadj.size();
  d.kU = (const uint64_t*)(uintptr_t)kU.size();
  d.kV = (const uint64_t*)(uintptr_t)kV.size();
  d.nU = g.nU;
  d.nV = g.nV;
  uint32_t dmax = 0;
  for (uint32_t v = 0; v < g.nV; ++v)
    dmax = std::max<uint32_t>(dmax, (uint32_t)(g.v_off[v + 1] - g.v_off[v]));
  d.dmaxV = dmax;
  d.swapped = (uint32_t)g.swapped;
  d.bound_height = bh;
  d.bound_size = bs;
  voff.insert(voff.end(), g.v_off.begin(), g.v_off.end());
  uoff.insert(uoff.end(), g.u_off.begin(), g.u_off.end());
  vadj.insert(vadj.end(), g.v_adj.begin(), g.v_adj.end());
  uadj.insert(uadj.end(), g.u_adj.begin(), g.u_adj.end());
  std::vector<uint64_t> ku, kv;
  hash_keys(g, ku, kv);
  kU.insert(kU.end(), ku.begin(), ku.end());
  kV.insert(kV.end(), kv.begin(), kv.end());
  table.push_back(d);
  double c = 0;
  for (uint32_t u = 0; u < g.nU; ++u) {
    const double du = (double)(g.u_off[u + 1] - g.u_off[u]);
    c += du * du;
  }
  cost.push_back(c);
  max_nV = std::max(max_nV, g.nV);
  max_dmaxV = std::max(max_dmaxV, dmax);
  return id;
}

void LanePack::append(const LanePack& o) {
  const uintptr_t bvo = voff.size(), bva = vadj.size(), buo = uoff.size(), bua = uadj.size(),
                  bku = kU.size(), bkv = kV.size();
  for (mbe::DevGraph d : o.table) {
    d.v_off = (const int64_t*)((uintptr_t)d.v_off + bvo);
    d.v_adj = (const uint32_t*)((uintptr_t)d.v_adj + bva);
    d.u_off = (const int64_t*)((uintptr_t)d.u_off + buo);
    d.u_adj = (const uint32_t*)((uintptr_t)d.u_adj + bua);
    d.kU = (const uint64_t*)((uintptr_t)d.kU + bku);
    d.kV = (const uint64_t*)((uintptr_t)d.kV + bkv);
    table.push_back(d);
  }
  cost.insert(cost.end(), o.cost.begin(), o.cost.end());
  voff.insert(voff.end(), o.voff.begin(), o.voff.end());
  uoff.insert(uoff.end(), o.uoff.begin(), o.uoff.end());
  vadj.insert(vadj.end(), o.vadj.begin(), o.vadj.end());
  uadj.insert(uadj.end(), o.uadj.begin(), o.uadj.end());
  kU.insert(kU.end(), o.kU.begin(), o.kU.end());
  kV.insert(kV.end(), o.kV.begin(), o.kV.end());
  max_nV = std::max(max_nV, o.max_nV);
  max_dmaxV = std::max(max_dmaxV, o.max_dmaxV);
}

void LanePack::finish(bool light_first) {
  const uint32_t n = (uint32_t)table.size();
  order.resize(n);
  for (uint32_t i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    return light_first ? cost[a] < cost[b] : cost[a] > cost[b];
  });
  root_base.assign(n + 1, 0ull);
  for (uint32_t j = 0; j < n; ++j) root_base[j + 1] = root_base[j] + table[order[j]].nV;
}

void LanePack::finish_classes(const std::vector<int>& cls, int n_classes) {
  const uint32_t n = (uint32_t)table.size();
  order.resize(n);
  for (uint32_t i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    return cls[a] != cls[b] ? cls[a] < cls[b] : cost[a] > cost[b];
  });
  seg_first.clear();
  seg_class.clear();
  for (uint32_t j = 0; j < n; ++j)
    if (j == 0 || cls[order[j]] != cls[order[j - 1]]) {
      seg_first.push_back(j);
      seg_class.push_back(cls[order[j]] < n_classes ? cls[order[j]] : n_classes - 1);
    }
  seg_first.push_back(n);
  // one root_base segment per class: its n_c + 1 prefix sums, from 0
  root_base.clear();
  for (size_t s = 0; s + 1 < seg_first.size(); ++s) {
    uint64_t acc = 0;
    root_base.push_back(0ull);
    for (uint32_t j = seg_first[s]; j < seg_first[s + 1]; ++j) {
      acc += table[order[j]].nV;
      root_base.push_back(acc);
    }
  }
  if (n == 0) root_base.assign(1, 0ull);
}

uint64_t LanePack::total_roots() const {
  if (root_base.empty()) return 0;
  if (seg_first.empty()) return root_base.back();
  uint64_t t = 0;
  for (size_t s = 0; s + 1 < seg_first.size(); ++s) t += root_base[seg_first[s + 1] + s];
  return t;
}

void LanePack::clear() {
  voff.clear();
  uoff.clear();
  vadj.clear();
  uadj.clear();
  kU.clear();
  kV.clear();
  table.clear();
  cost.clear();
  order.clear();
  root_base.clear();
  seg_first.clear();
  seg_class.clear();
  max_nV = max_dmaxV = 0;
}

// ---------------------------------------------------------------------------
// Lane

int Lane::alloc(Engine* e, const LaneCaps& c, const LaneKnobs& k) {
  eng = e;
  caps = c;
  knobs = k;
  LANE_TRY(hipSetDevice(eng->device));
  LANE_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  LANE_TRY(hipEventCreate(&ev0));
  LANE_TRY(hipEventCreate(&ev1));
  const uint64_t ng = std::max<uint32_t>(c.max_graphs, 1);
  LANE_TRY(hipMalloc(&voff, std::max<uint64_t>(c.max_voff, 1) * 8));
  LANE_TRY(hipMalloc(&uoff, std::max<uint64_t>(c.max_uoff, 1) * 8));
  LANE_TRY(hipMalloc(&vadj, std::max<uint64_t>(c.max_edges, 1) * 4));
  LANE_TRY(hipMalloc(&uadj, std::max<uint64_t>(c.max_edges, 1) * 4));
  LANE_TRY(hipMalloc(&kU, std::max<uint64_t>(c.max_uoff, 1) * 8));
  LANE_TRY(hipMalloc(&kV, std::max<uint64_t>(c.max_voff, 1) * 8));
  LANE_TRY(hipMalloc(&gt, ng * sizeof(mbe::DevGraph)));
  LANE_TRY(hipMalloc(&root_base, (ng + kLaneClasses) * 8));
  LANE_TRY(hipMalloc(&order, ng * 4));
  LANE_TRY(hipMalloc(&per_graph, ng * kPerGraph * 8));
  LANE_TRY(hipMalloc(&ctl, kCtlWords * 8));
  LANE_TRY(hipMalloc(&acc, mbe::A_COUNT * 8));
  LANE_TRY(hipMalloc(&deferred, knobs.deferred_cap * sizeof(mbe::Task)));
  // the arrays of a full lane plus the 16-byte alignment of each of the nine
  // pieces an upload stages
  stage_bytes = (c.max_voff + c.max_uoff) * 16 + c.max_edges * 8 +
                ng * (sizeof(mbe::DevGraph) + 8 + 4) + 8 * kLaneClasses + 9 * 16;
  LANE_TRY(hipHostMalloc((void**)&stage, stage_bytes, hipHostMallocDefault));
  LANE_TRY(hipHostMalloc((void**)&out_host, (ng * kPerGraph + mbe::A_COUNT) * 8,
                         hipHostMallocDefault));
  return alloc_run();
}

KernelChoice Lane::class_kernel(int c, int* blocks) const {
  const int w = c >= 0 && c < kLaneClasses && cls_waves[c] > 0 ? cls_waves[c] : knobs.waves;
  const int l = c >= 0 && c < kLaneClasses && cls_lds[c] >= 0 ? cls_lds[c] : knobs.lds_kb;
  const KernelChoice k = pick_kernel(l, w, false);
  if (blocks) *blocks = k.fn ? (knobs.blocks > 0 ? knobs.blocks : resident_blocks(k, eng->cu_count)) : 0;
  return k;
}

int Lane::alloc_run() {
  kc = pick_kernel(knobs.lds_kb, knobs.waves, false);
  wk = pick_kernel(8, 4, true);
  if (!kc.fn)
    return set_error(GMBE_ERR_ARG, "no kernel with %d waves per block and lds_kb=%d",
                     knobs.waves, knobs.lds_kb);
  const int blocks = knobs.blocks > 0 ? knobs.blocks : resident_blocks(kc, eng->cu_count);
  // the wide pass runs on at most as many waves as the main pass
  wblocks = std::max(1, std::min(resident_blocks(wk, eng->cu_count), blocks * kc.waves / wk.waves));
  // slabs for the largest grid of any class's kernel
  int nw = blocks * kc.waves;
  for (int c = 0; c < kLaneClasses; ++c) {
    int cb = 0;
    const KernelChoice k = class_kernel(c, &cb);
    if (!k.fn)
      return set_error(GMBE_ERR_ARG, "no kernel for class %d (waves per block %d, lds_kb %d)", c,
                       cls_waves[c], cls_lds[c]);
    nw = std::max(nw, cb * k.waves);
  }
  mbe::Scratch lay = scratch_layout(std::max<uint32_t>(caps.max_nV, 1),
                                    std::max<uint32_t>(caps.max_dmaxV, 1),
                                    (uint64_t)knobs.arena_kb * 1024ull,
                                    (uint64_t)knobs.ctx_kb * 1024ull, (uint32_t)knobs.ctx_slots);
  const uint64_t bytes = lay.slab * (uint64_t)nw;
  if (bytes != scratch_bytes) {
    if (scratch) LANE_TRY(hipFree(scratch));
    scratch = nullptr;
    scratch_bytes = 0;
    hipError_t e = hipMalloc(&scratch, bytes);
    if (e != hipSuccess)
      return set_error(GMBE_ERR_NOMEM, "lane scratch of %llu bytes (%d waves): %s",
                       (unsigned long long)bytes, nw, hipGetErrorString(e));
    scratch_bytes = bytes;
    // slabs start zeroed (the candidate maps rely on it); a kernel, not a
    // memset, so that nothing here differs between eager and captured use
    LANE_TRY(launch_fill(st, (unsigned long long*)scratch, bytes / 8, 0ull));
  }
  lay.base = scratch;
  layout = lay;
  nwaves = nw;
  nblocks = blocks;
  uint64_t qc = (1ull << knobs.queue_log2) / mbe::kShards;
  if (qc < 16) qc = 16;
  if (qc != qcap) {
    if (q) LANE_TRY(hipFree(q));
    if (qseq) LANE_TRY(hipFree(qseq));
    q = nullptr;
    qseq = nullptr;
    LANE_TRY(hipMalloc(&q, qc * mbe::kShards * mbe::kEntryWords * 4));
    LANE_TRY(hipMalloc(&qseq, qc * mbe::kShards * 8));
    qcap = qc;
  }
  LANE_TRY(hipStreamSynchronize(st));
  return GMBE_OK;
}

void Lane::release() {
  if (eng && eng->device >= 0) (void)hipSetDevice(eng->device);
  if (st) (void)hipStreamSynchronize(st);
  void* dev[] = {voff, uoff, vadj, uadj, kU, kV, gt, root_base, order, per_graph,
                 scratch, q, qseq, ctl, acc, deferred};
  for (void* p : dev)
    if (p) (void)hipFree(p);
  if (stage) (void)hipHostFree(stage);
  if (out_host) (void)hipHostFree(out_host);
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  if (st) (void)hipStreamDestroy(st);
  *this = Lane();
}

int Lane::upload(const LanePack& p) {
  const uint32_t n = (uint32_t)p.table.size();
  if (n > caps.max_graphs || p.voff.size() > caps.max_voff || p.uoff.size() > caps.max_uoff ||
      p.vadj.size() > caps.max_edges || p.max_nV > caps.max_nV || p.max_dmaxV > caps.max_dmaxV)
    return set_error(GMBE_ERR_CAPACITY, "graph set exceeds the lane's capacity");
  // stage: voff | uoff | kU | kV | vadj | uadj | table | root_base | order
  char* s = stage;
  auto put = [&](const void* src, uint64_t bytes, void* dst) -> hipError_t {
    if (bytes == 0) return hipSuccess;
    memcpy(s, src, bytes);
    hipError_t e = hipMemcpyAsync(dst, s, bytes, hipMemcpyHostToDevice, st);
    s += (bytes + 15) / 16 * 16;
    return e;
  };
  std::vector<mbe::DevGraph> tab(p.table);
  for (mbe::DevGraph& d : tab) {
    d.v_off = voff + (uintptr_t)d.v_off;
    d.v_adj = vadj + (uintptr_t)d.v_adj;
    d.u_off = uoff + (uintptr_t)d.u_off;
    d.u_adj = uadj + (uintptr_t)d.u_adj;
    d.kU = kU + (uintptr_t)d.kU;
    d.kV = kV + (uintptr_t)d.kV;
  }
  LANE_TRY(put(p.voff.data(), p.voff.size() * 8, voff));
  LANE_TRY(put(p.uoff.data(), p.uoff.size() * 8, uoff));
  LANE_TRY(put(p.kU.data(), p.kU.size() * 8, kU));
  LANE_TRY(put(p.kV.data(), p.kV.size() * 8, kV));
  LANE_TRY(put(p.vadj.data(), p.vadj.size() * 4, vadj));
  LANE_TRY(put(p.uadj.data(), p.uadj.size() * 4, uadj));
  LANE_TRY(put(tab.data(), tab.size() * sizeof(mbe::DevGraph), gt));
  LANE_TRY(put(p.root_base.data(), p.root_base.size() * 8, root_base));
  LANE_TRY(put(p.order.data(), p.order.size() * 4, order));
  n_graphs = n;
  n_roots = p.total_roots();
  segs.clear();
  for (size_t s = 0; s + 1 < p.seg_first.size(); ++s) {
    LaneSeg g;
    g.g0 = p.seg_first[s];
    g.n = p.seg_first[s + 1] - p.seg_first[s];
    g.rb_off = g.g0 + s;
    g.roots = p.root_base[g.rb_off + g.n];
    g.cls = p.seg_class[s];
    segs.push_back(g);
  }
  return GMBE_OK;
}

int Lane::upload_many(const std::vector<const LanePack*>& parts, bool light_first) {
  uint64_t nvo = 0, nuo = 0, ne = 0;
  uint32_t n = 0, mv = 0, md = 0;
  for (const LanePack* p : parts) {
    nvo += p->voff.size();
    nuo += p->uoff.size();
    ne += p->vadj.size();
    n += (uint32_t)p->table.size();
    mv = std::max(mv, p->max_nV);
    md = std::max(md, p->max_dmaxV);
  }
  if (n > caps.max_graphs || nvo > caps.max_voff || nuo > caps.max_uoff || ne > caps.max_edges ||
      mv > caps.max_nV || md > caps.max_dmaxV)
    return set_error(GMBE_ERR_CAPACITY, "graph set exceeds the lane's capacity");
  // stage layout: voff | uoff | kU | kV | vadj | uadj | table | root_base | order
  auto al = [](uint64_t b) { return (b + 15) / 16 * 16; };
  char* s_voff = stage;
  char* s_uoff = s_voff + al(nvo * 8);
  char* s_kU = s_uoff + al(nuo * 8);
  char* s_kV = s_kU + al(nuo * 8);
  char* s_vadj = s_kV + al(nvo * 8);
  char* s_uadj = s_vadj + al(ne * 4);
  char* s_tab = s_uadj + al(ne * 4);
  char* s_rb = s_tab + al((uint64_t)n * sizeof(mbe::DevGraph));
  char* s_ord = s_rb + al(((uint64_t)n + 1) * 8);
  mbe::DevGraph* tab = (mbe::DevGraph*)s_tab;
  std::vector<double> cost;
  cost.reserve(n);
  uint64_t bvo = 0, buo = 0, bva = 0, bua = 0, bku = 0, bkv = 0;
  uint32_t gi = 0;
  auto cp = [](char* dst, const void* src, uint64_t bytes) {
    if (bytes) memcpy(dst, src, bytes);
  };
  for (const LanePack* p : parts) {
    for (size_t j = 0; j < p->table.size(); ++j) {
      mbe::DevGraph d = p->table[j];
      d.v_off = voff + bvo + (uintptr_t)d.v_off;
      d.v_adj = vadj + bva + (uintptr_t)d.v_adj;
      d.u_off = uoff + buo + (uintptr_t)d.u_off;
      d.u_adj = uadj + bua + (uintptr_t)d.u_adj;
      d.kU = kU + bku + (uintptr_t)d.kU;
      d.kV = kV + bkv + (uintptr_t)d.kV;
      tab[gi++] = d;
      cost.push_back(p->cost[j]);
    }
    cp(s_voff + bvo * 8, p->voff.data(), p->voff.size() * 8);
    cp(s_uoff + buo * 8, p->uoff.data(), p->uoff.size() * 8);
    cp(s_kU + bku * 8, p->kU.data(), p->kU.size() * 8);
    cp(s_kV + bkv * 8, p->kV.data(), p->kV.size() * 8);
    cp(s_vadj + bva * 4, p->vadj.data(), p->vadj.size() * 4);
    cp(s_uadj + bua * 4, p->uadj.data(), p->uadj.size() * 4);
    bvo += p->voff.size();
    buo += p->uoff.size();
    bku += p->kU.size();
    bkv += p->kV.size();
    bva += p->vadj.size();
    bua += p->uadj.size();
  }
  uint32_t* ord = (uint32_t*)s_ord;
  for (uint32_t i = 0; i < n; ++i) ord[i] = i;
  std::stable_sort(ord, ord + n, [&](uint32_t a, uint32_t b) {
    return light_first ? cost[a] < cost[b] : cost[a] > cost[b];
  });
  unsigned long long* rb = (unsigned long long*)s_rb;
  rb[0] = 0;
  for (uint32_t j = 0; j < n; ++j) rb[j + 1] = rb[j] + tab[ord[j]].nV;
  auto h2d = [&](void* dst, const char* src, uint64_t bytes) -> hipError_t {
    return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st) : hipSuccess;
  };
  LANE_TRY(h2d(voff, s_voff, nvo * 8));
  LANE_TRY(h2d(uoff, s_uoff, nuo * 8));
  LANE_TRY(h2d(kU, s_kU, nuo * 8));
  LANE_TRY(h2d(kV, s_kV, nvo * 8));
  LANE_TRY(h2d(vadj, s_vadj, ne * 4));
  LANE_TRY(h2d(uadj, s_uadj, ne * 4));
  LANE_TRY(h2d(gt, s_tab, (uint64_t)n * sizeof(mbe::DevGraph)));
  LANE_TRY(h2d(root_base, s_rb, ((uint64_t)n + 1) * 8));
  LANE_TRY(h2d(order, s_ord, (uint64_t)n * 4));
  n_graphs = n;
  n_roots = rb[n];
  segs.clear();
  return GMBE_OK;
}

mbe::Params Lane::params(bool wide) const {
  mbe::Params P{};
  P.gt = gt;
  P.per_graph = per_graph;
  P.s.root_next = ctl + C_ROOT_NEXT * kCtlStride;
  P.s.work_left = ctl + C_WORK_LEFT * kCtlStride;
  P.s.qctl = ctl + C_QCTL * kCtlStride;
  P.s.queued = ctl + C_QUEUED * kCtlStride;
  P.list_cursor = ctl + C_LIST_CURSOR * kCtlStride;
  P.s.n_deferred = ctl + C_DEFERRED * kCtlStride;
  P.s.deadline_word = ctl + C_DEADLINE * kCtlStride;  // one absolute deadline for both passes
  P.s.q = q;
  P.s.qseq = qseq;
  P.s.qcap = qcap;
  P.s.stride = 1;
  P.s.phase = 0;
  P.s.n_graphs = n_graphs;
  if (!wide) {
    P.s.n_roots = n_roots;
    P.s.root_base = root_base;
    P.s.order = order;
    P.s.deferred = deferred;
    P.s.deferred_cap = knobs.deferred_cap;
  } else {
    P.s.n_roots_dev = ctl + C_WIDE_N * kCtlStride;
    P.s.root_tasks = deferred;
  }
  P.w = layout;
  P.acc = acc;
  P.list = nullptr;
  P.list_cap = 0;
  P.prune = knobs.prune;
  P.max_prefix = knobs.max_prefix;
  P.split = knobs.split;
  P.profile = 0;
  P.group = knobs.group;
  P.share = knobs.share && knobs.split;
  P.merge_min = (int)knobs.merge_min;
  P.merge_branch = (int)knobs.merge_branch;
  P.share_work = (int)std::min<int64_t>(knobs.share_work, INT32_MAX);
  P.share_queue = knobs.share_queue >= 0 ? (int)knobs.share_queue : nwaves;
  P.late_share = P.share ? (int)knobs.late_share : 0;
  P.dom_k = knobs.dom_k;
  P.dom_skip = knobs.dom_skip;
  P.recompute = knobs.recompute;
  P.count_bytes = knobs.count_bytes;
  P.budget = (uint64_t)knobs.budget_ms * (uint64_t)eng->wall_khz;
  return P;
}

int Lane::record_run() {
  const mbe::Params P = params(false), Pw = params(true);
  const uint64_t ng = std::max<uint32_t>(n_graphs, 1);
  LANE_TRY(launch_fill(st, acc, mbe::A_COUNT, 0ull));
  LANE_TRY(launch_fill(st, per_graph, ng * kPerGraph, 0ull));
  if (n_graphs > 0) LANE_TRY(launch_init_left(st, P, n_graphs));
  if (segs.empty()) {
    LANE_TRY(launch_ctl_start(st, ctl, n_roots));
    LANE_TRY(launch_init_queue(st, qseq, qcap));
    LANE_TRY(launch_reset_ctx(st, P, (unsigned long long)nwaves * layout.ctx_slots));
    if (n_roots > 0) LANE_TRY(launch_persistent(kc, nblocks, st, P));
    // wide pass over the deferred tasks; its grid exits at once when there are none
    LANE_TRY(launch_ctl_wide(st, ctl, knobs.deferred_cap));
    LANE_TRY(launch_init_queue(st, qseq, qcap));
    if (n_roots > 0) LANE_TRY(launch_persistent(wk, wblocks, st, Pw));
  } else {
    // one main and one wide launch per size class, each over the class's
    // segment of the claim order with the class's kernel variant
    bool first = true;
    for (const LaneSeg& g : segs) {
      if (g.roots == 0) continue;
      int cb = 0;
      const KernelChoice ck = class_kernel(g.cls, &cb);
      mbe::Params Pc = P;
      Pc.s.n_roots = g.roots;
      Pc.s.root_base = root_base + g.rb_off;
      Pc.s.order = order + g.g0;
      Pc.s.n_graphs = g.n;
      if (knobs.share_queue < 0) Pc.share_queue = cb * ck.waves;
      const int cwb = std::max(1, std::min(resident_blocks(wk, eng->cu_count), cb * ck.waves / wk.waves));
      LANE_TRY(launch_ctl_start(st, ctl, g.roots, !first));  // one deadline for the run
      first = false;
      LANE_TRY(launch_init_queue(st, qseq, qcap));
      LANE_TRY(launch_reset_ctx(st, Pc, (unsigned long long)nwaves * layout.ctx_slots));
      LANE_TRY(launch_persistent(ck, cb, st, Pc));
      LANE_TRY(launch_ctl_wide(st, ctl, knobs.deferred_cap));
      LANE_TRY(launch_init_queue(st, qseq, qcap));
      LANE_TRY(launch_persistent(wk, cwb, st, Pw));
    }
  }
  LANE_TRY(hipMemcpyAsync(out_host, per_graph, ng * kPerGraph * 8, hipMemcpyDeviceToHost, st));
  LANE_TRY(hipMemcpyAsync(out_host + ng * kPerGraph, acc, mbe::A_COUNT * 8,
                          hipMemcpyDeviceToHost, st));
  return GMBE_OK;
}

void Lane::result(uint32_t i, gmbe_result* r) const {
  memset(r, 0, sizeof(*r));
  const unsigned long long* a = out_host + (uint64_t)i * kPerGraph;
  r->maximal = a[mbe::G_MAXIMAL];
  r->hash_sum = a[mbe::G_HSUM];
  r->hash_xor = a[mbe::G_HXOR];
  r->non_maximal = a[mbe::G_NONMAX];
  r->nodes = a[mbe::G_NODES];
  r->overflow = a[mbe::G_OVERFLOW];
  r->alg_bytes = a[mbe::G_BALG];
  r->tasks = a[mbe::G_TASKS];
}

bool Lane::complete(uint32_t i) const {
  return out_host[(uint64_t)i * kPerGraph + mbe::G_LEFT] == 0;
}

void Lane::totals(gmbe_result* r) const {
  memset(r, 0, sizeof(*r));
  for (uint32_t i = 0; i < n_graphs; ++i) {
    gmbe_result g;
    result(i, &g);
    r->maximal += g.maximal;
    r->hash_sum += g.hash_sum;
    r->hash_xor ^= g.hash_xor;
    r->non_maximal += g.non_maximal;
    r->nodes += g.nodes;
    r->overflow += g.overflow;
    r->alg_bytes += g.alg_bytes;
    r->tasks += g.tasks;
  }
  const unsigned long long* a = out_host + (uint64_t)std::max<uint32_t>(n_graphs, 1) * kPerGraph;
  r->splits = a[mbe::A_SPLITS];
}

}  // namespace gmbe
