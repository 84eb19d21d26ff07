// finder.hip — persistent enumeration kernel and the plain finder behind the
// C ABI (gmbe_finder_*).  Replaces IterFinderGpu6's constructor / Execute /
// destructor (src/IterFinderGpu.cu:1056-1119) and the kernel launch
// IterFinderKernel_6<<<MAX_BLOCKS, 512>>> (src/IterFinderGpu.cu:1104-1107).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "device_graph.hpp"
#include "graph.hpp"
#include "mbe_kernel.hpp"

using namespace gmbe;

namespace mbe {

// Mark every shared-context slot free (a run cut short by its deadline can
// leave items of a context unrun).
__global__ void reset_ctx(Params P, unsigned long long n) {
  for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (unsigned long long)gridDim.x * blockDim.x)
    ctx_hdr(P, i)->refs = 0ull;
}

// Outstanding units of every graph of a batch: its top-level vertices.
__global__ void init_left(Params P, unsigned long long n) {
  for (unsigned long long g = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; g < n;
       g += (unsigned long long)gridDim.x * blockDim.x)
    P.per_graph[g * kPerGraph + G_LEFT] = P.gt[g].nV;
}

// Reset the queue's slot sequence numbers (slot i of every shard starts free
// for position i).
__global__ void init_queue(unsigned long long* seq, unsigned long long cap) {
  const unsigned long long total = cap * kShards;
  for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
       i < total; i += (unsigned long long)gridDim.x * blockDim.x)
    seq[i] = i & (cap - 1);
}

// Fill n u64 words with v.
__global__ void fill_u64(unsigned long long* p, unsigned long long n, unsigned long long v) {
  for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (unsigned long long)gridDim.x * blockDim.x)
    p[i] = v;
}

// Control block of a launch: everything zero but work_left.
__global__ void ctl_start(unsigned long long* ctl, unsigned long long words, unsigned long long wl,
                          unsigned long long n_roots, unsigned long long keep) {
  const unsigned long long kept = keep < words ? ctl[keep] : 0ull;
  __syncthreads();
  for (unsigned long long i = threadIdx.x; i < words; i += blockDim.x)
    ctl[i] = i == wl ? n_roots : i == keep ? kept : 0ull;
}

// Control block of the wide pass, from the deferred count the main pass left.
// A main pass cut short by its deadline (work left) gets no wide pass: its
// deferred tasks stay unfinished and their graphs incomplete.
__global__ void ctl_wide(unsigned long long* ctl, unsigned long long words, unsigned long long wl,
                         unsigned long long nd, unsigned long long wn, unsigned long long cur,
                         unsigned long long dlw, unsigned long long cap) {
  const unsigned long long n = ctl[wl] != 0 ? 0ull : ctl[nd] < cap ? ctl[nd] : cap;
  const unsigned long long cursor = ctl[cur], deadline = ctl[dlw];
  __syncthreads();
  for (unsigned long long i = threadIdx.x; i < words; i += blockDim.x) {
    unsigned long long v = 0ull;
    if (i == wl || i == wn) v = n;
    if (i == cur) v = cursor;
    if (i == dlw) v = deadline;  // the main pass's deadline holds for the wide pass
    ctl[i] = v;
  }
}

}  // namespace mbe

namespace gmbe {

#define HIP_TRY(expr)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return set_error(GMBE_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                       __FILE__, __LINE__);                                                  \
  } while (0)

using mbe::kCtlStride;
using mbe::kShards;

KernelChoice pick_kernel(int lds_kb, int waves, bool wide) {
  // four waves per SIMD where the LDS allows (a 7 KB list arena per wave
  // leaves room for four blocks of four waves per CU); the wide-bitset
  // kernel needs two.  Each variant is compiled in its own kern_*.hip.
  if (wide) return kernel_wide();
  if (waves == 8 && lds_kb == 7) return kernel_w8_l7();
  if (waves == 2 && lds_kb == 7) return kernel_w2_l7();
  if (waves != 4) return {nullptr, 0, 0};
  switch (lds_kb) {
    case 7: return kernel_w4_l7();
    case 0: return kernel_w4_l0();
    case 3: return kernel_w4_l3();
    case 16: return kernel_w4_l16();
    default: return {nullptr, 0, 0};
  }
}

int resident_blocks(const KernelChoice& kc, int cu_count) {
  int per_sm = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_sm, (const void*)kc.fn, kc.waves * 64,
                                                   0) != hipSuccess ||
      per_sm < 1)
    per_sm = 1;
  return per_sm * cu_count;
}

hipError_t launch_persistent(const KernelChoice& kc, int blocks, hipStream_t st,
                             const mbe::Params& P) {
  hipLaunchKernelGGL(kc.fn, dim3(blocks), dim3(kc.waves * 64), 0, st, P);
  return hipGetLastError();
}
hipError_t launch_init_queue(hipStream_t st, unsigned long long* seq, unsigned long long cap) {
  hipLaunchKernelGGL(mbe::init_queue, dim3(256), dim3(256), 0, st, seq, cap);
  return hipGetLastError();
}
hipError_t launch_init_left(hipStream_t st, const mbe::Params& P, unsigned long long n) {
  hipLaunchKernelGGL(mbe::init_left, dim3(64), dim3(256), 0, st, P, n);
  return hipGetLastError();
}
hipError_t launch_reset_ctx(hipStream_t st, const mbe::Params& P, unsigned long long n) {
  hipLaunchKernelGGL(mbe::reset_ctx, dim3(64), dim3(256), 0, st, P, n);
  return hipGetLastError();
}
hipError_t launch_fill(hipStream_t st, unsigned long long* p, unsigned long long n,
                       unsigned long long v) {
  if (n == 0) return hipSuccess;
  const unsigned long long blocks = std::min<unsigned long long>((n + 255) / 256, 1024);
  hipLaunchKernelGGL(mbe::fill_u64, dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}
hipError_t launch_ctl_start(hipStream_t st, unsigned long long* ctl, unsigned long long n_roots,
                            bool keep_deadline) {
  hipLaunchKernelGGL(mbe::ctl_start, dim3(1), dim3(256), 0, st, ctl, kCtlWords,
                     (unsigned long long)C_WORK_LEFT * kCtlStride, n_roots,
                     keep_deadline ? (unsigned long long)C_DEADLINE * kCtlStride : ~0ull);
  return hipGetLastError();
}
hipError_t launch_ctl_wide(hipStream_t st, unsigned long long* ctl, unsigned long long cap) {
  hipLaunchKernelGGL(mbe::ctl_wide, dim3(1), dim3(256), 0, st, ctl, kCtlWords,
                     (unsigned long long)C_WORK_LEFT * kCtlStride,
                     (unsigned long long)C_DEFERRED * kCtlStride,
                     (unsigned long long)C_WIDE_N * kCtlStride,
                     (unsigned long long)C_LIST_CURSOR * kCtlStride,
                     (unsigned long long)C_DEADLINE * kCtlStride, cap);
  return hipGetLastError();
}

int Engine::ensure_device(int dev) {
  if (device >= 0) return GMBE_OK;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0)
    return set_error(GMBE_ERR_HIP, "no HIP device available (%s)", hipGetErrorString(e));
  HIP_TRY(hipSetDevice(dev));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, dev));
  cu_count = prop.multiProcessorCount;
  int khz = 0;
  HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
  wall_khz = khz > 0 ? khz : 100000;
  device = dev;
  return GMBE_OK;
}

int DeviceGraph::upload(const gmbe_graph& g) {
  nU = g.nU;
  nV = g.nV;
  E = g.E;
  swapped = g.swapped;
  dmaxV = 0;
  for (uint32_t v = 0; v < g.nV; ++v)
    dmaxV = std::max<uint32_t>(dmaxV, (uint32_t)(g.v_off[v + 1] - g.v_off[v]));
  std::vector<uint64_t> kU, kV;
  hash_keys(g, kU, kV);
  size_t b_voff = (size_t)(nV + 1) * 8, b_uoff = (size_t)(nU + 1) * 8;
  size_t b_adj = (size_t)std::max<uint64_t>(E, 1) * 4;
  size_t b_kU = (size_t)std::max<uint32_t>(nU, 1) * 8, b_kV = (size_t)std::max<uint32_t>(nV, 1) * 8;
  HIP_TRY(hipMalloc(&v_off, b_voff));
  HIP_TRY(hipMalloc(&u_off, b_uoff));
  HIP_TRY(hipMalloc(&v_adj, b_adj));
  HIP_TRY(hipMalloc(&u_adj, b_adj));
  HIP_TRY(hipMalloc(&kUd, b_kU));
  HIP_TRY(hipMalloc(&kVd, b_kV));
  HIP_TRY(hipMemcpy(v_off, g.v_off.data(), b_voff, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(u_off, g.u_off.data(), b_uoff, hipMemcpyHostToDevice));
  if (E) {
    HIP_TRY(hipMemcpy(v_adj, g.v_adj.data(), (size_t)E * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(u_adj, g.u_adj.data(), (size_t)E * 4, hipMemcpyHostToDevice));
  }
  if (nU) HIP_TRY(hipMemcpy(kUd, kU.data(), (size_t)nU * 8, hipMemcpyHostToDevice));
  if (nV) HIP_TRY(hipMemcpy(kVd, kV.data(), (size_t)nV * 8, hipMemcpyHostToDevice));
  return GMBE_OK;
}

void DeviceGraph::release() {
  (void)hipFree(v_off);
  (void)hipFree(u_off);
  (void)hipFree(v_adj);
  (void)hipFree(u_adj);
  (void)hipFree(kUd);
  (void)hipFree(kVd);
  v_off = u_off = nullptr;
  v_adj = u_adj = nullptr;
  kUd = kVd = nullptr;
}

mbe::DevGraph DeviceGraph::view() const {
  mbe::DevGraph d;
  d.v_off = v_off;
  d.v_adj = v_adj;
  d.u_off = u_off;
  d.u_adj = u_adj;
  d.kU = kUd;
  d.kV = kVd;
  d.nU = nU;
  d.nV = nV;
  d.dmaxV = dmaxV;
  d.swapped = (uint32_t)swapped;
  return d;
}

static uint64_t align_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

// Per-wave slab layout for a graph with nV enumeration-side vertices and
// maximum enumeration-side degree dmaxV.
mbe::Scratch scratch_layout(uint32_t nV, uint32_t dmaxV, uint64_t arena_bytes,
                            uint64_t ctx_bytes, uint32_t ctx_slots) {
  mbe::Scratch s{};
  uint64_t o = 0;
  auto take = [&](uint64_t bytes) {
    uint64_t r = o;
    o = align_up(o + bytes, 256);
    return r;
  };
  uint64_t nv = std::max<uint32_t>(nV, 1), d = (uint64_t)dmaxV + 2;
  s.nW0 = (uint32_t)((nv + 63) / 64);
  s.nW1 = (uint32_t)((s.nW0 + 63) / 64);
  s.o_cmap = take(nv * 4);
  s.o_bm0 = take((uint64_t)s.nW0 * 8);
  s.o_bm1 = take((uint64_t)s.nW1 * 8);
  s.o_L = take(d * 4);
  s.o_Lp = take(d * 4);
  s.o_T = take(d * 4);
  s.o_kL = take(d * 8);
  s.o_U = take(nv * 4);
  s.o_Ucnt = take(nv * 4);
  s.o_C = take(nv * 4);
  s.o_kC = take(nv * 8);
  s.o_R = take(nv * 4);
  s.o_fr = take((d + 2) * sizeof(mbe::Frame));
  s.o_dom = take((uint64_t)mbe::kMaxDom * mbe::kMaxW * 8);
  s.o_arena = take(arena_bytes);
  s.arena_bytes = arena_bytes;
  s.o_ctx = take(ctx_bytes * ctx_slots);
  s.ctx_bytes = ctx_bytes;
  s.ctx_slots = ctx_slots;
  s.slab = o;
  return s;
}

}  // namespace gmbe

// ---------------------------------------------------------------------------

struct gmbe_finder {
  const gmbe_graph* g = nullptr;
  int sel = 2, bound_height = 20, bound_size = 1500;
  // knobs
  int64_t waves_per_block = 4, blocks = 0, prune = 1, max_prefix = 4, lds_kb = 7;
  int64_t queue_log2 = 22, list_cap = 0, arena_kb = 16384, root_stride = 1, root_phase = 0;
  int64_t root_begin = 0, root_end = -1;  // processing-order rank range, -1 = all
  int64_t deadline_ms = 0;                 // 0 = run to completion
  int64_t late_share = 32;                 // 0 = off
  int64_t dom_k = 8, dom_skip = 1;         // list reduction (reduce_list)
  int64_t recompute = 0;                   // split large tasks without a free context into children
  int64_t count_bytes = 0;                 // count algorithmic bytes (alg_bytes)
  int64_t reduce_min = 0;                  // reduce only lists at least this long
  int64_t wide_resident = 1;               // wide pass: launch only the blocks that fit at once
  int64_t ctx_slots = 16;                  // shared-context slots per wave
  int64_t profile = 0, group = 1, share = 1, ctx_kb = 4096, share_work = 1024, share_queue = -1;
  int64_t merge_min = 256, merge_branch = 16;
  bool arena_set = false, ctx_set = false;  // sizes given by the caller
  unsigned long long prof[mbe::A_COUNT] = {};
  gmbe::Engine eng;
  gmbe::DeviceGraph dg;
  // device buffers
  char* scratch = nullptr;
  uint64_t scratch_bytes = 0;
  mbe::Scratch layout{};
  int nwaves = 0, nblocks = 0;
  uint32_t* q = nullptr;
  unsigned long long* qseq = nullptr;
  uint64_t qcap = 0;
  // control counters, each on its own 256-byte line (kCtlStride u64 apart):
  // root_next, work_left, queued, list_cursor, n_deferred, then per-shard head/tail
  unsigned long long* ctl = nullptr;
  unsigned long long* acc = nullptr;
  uint32_t* list = nullptr;
  uint64_t list_words = 0;
  mbe::Task* deferred = nullptr;
  mbe::DevGraph* gt = nullptr;  // device graph table (one entry)
  int64_t exit_stamps = 0;              // knob: record per-block exit times
  unsigned long long* stamps = nullptr;  // 1 + nblocks: launch stamp, then per block
  std::vector<double> exit_s;            // per-block exit, seconds after launch
  uint64_t deferred_cap = 1ull << 20;
  uint64_t last_deferred = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, evm = nullptr;  // evm: end of the main kernel
  double main_ms = 0, wide_ms = 0;  // last execute: main kernel, wide-bitset pass
  unsigned long long main_acc[mbe::A_COUNT] = {};  // counters after the main kernel
  bool ran = false;
  uint64_t list_used = 0, list_ovf = 0;
};

using namespace gmbe;

static int finder_alloc(gmbe_finder* f) {
  KernelChoice kc = pick_kernel((int)f->lds_kb, (int)f->waves_per_block, false);
  int blocks = f->blocks > 0 ? (int)f->blocks : resident_blocks(kc, f->eng.cu_count);
  int nwaves = blocks * kc.waves;
  // unless set, the list arena and the context slots are sized for the graph:
  // a list holds at most |V| entries of <= 60 bytes
  const uint64_t nv = f->dg.nV;
  const int64_t arena_kb = f->arena_set ? f->arena_kb
                                        : std::min<int64_t>(16384, std::max<int64_t>(4096, (int64_t)(nv * 24 * 64 / 1024)));
  const int64_t ctx_kb = f->ctx_set ? f->ctx_kb
                                    : std::min<int64_t>(1024, std::max<int64_t>(256, (int64_t)(nv * 16 / 1024)));
  f->arena_kb = arena_kb;
  f->ctx_kb = ctx_kb;
  mbe::Scratch lay = scratch_layout(f->dg.nV, f->dg.dmaxV, (uint64_t)arena_kb * 1024ull,
                                    (uint64_t)ctx_kb * 1024ull, (uint32_t)f->ctx_slots);
  uint64_t bytes = lay.slab * (uint64_t)nwaves;
  if (bytes != f->scratch_bytes || nwaves != f->nwaves) {
    if (f->scratch) (void)hipFree(f->scratch);
    f->scratch = nullptr;
    f->scratch_bytes = 0;
    hipError_t e = hipMalloc(&f->scratch, bytes);
    if (e != hipSuccess)
      return set_error(GMBE_ERR_NOMEM, "scratch of %llu bytes (%d waves): %s",
                       (unsigned long long)bytes, nwaves, hipGetErrorString(e));
    HIP_TRY(hipMemsetAsync(f->scratch, 0, bytes, f->stream));
    f->scratch_bytes = bytes;
  }
  lay.base = f->scratch;
  f->layout = lay;
  f->nwaves = nwaves;
  f->nblocks = blocks;
  uint64_t qcap = (1ull << f->queue_log2) / kShards;  // slots per shard (a power of two)
  // at least two: a slot's sequence number tells "holds position p" (p + 1)
  // from "free for position p + qcap" (p + qcap) only when qcap >= 2
  if (qcap < 2) qcap = 2;
  if (qcap != f->qcap) {
    if (f->q) (void)hipFree(f->q);
    if (f->qseq) (void)hipFree(f->qseq);
    f->q = nullptr;
    f->qseq = nullptr;
    HIP_TRY(hipMalloc(&f->q, qcap * kShards * mbe::kEntryWords * sizeof(uint32_t)));
    HIP_TRY(hipMalloc(&f->qseq, qcap * kShards * 8));
    f->qcap = qcap;
  }
  if (!f->deferred) HIP_TRY(hipMalloc(&f->deferred, f->deferred_cap * sizeof(mbe::Task)));
  if ((uint64_t)f->list_cap != f->list_words) {
    if (f->list) (void)hipFree(f->list);
    f->list = nullptr;
    if (f->list_cap > 0) HIP_TRY(hipMalloc(&f->list, (size_t)f->list_cap * 4));
    f->list_words = (uint64_t)f->list_cap;
  }
  return GMBE_OK;
}

extern "C" {

static int gmbe_finder_create_impl(const gmbe_graph* g, int sel, int bound_height, int bound_size,
                                   int ngpus, gmbe_finder** out) {
  if (!g || !out) return set_error(GMBE_ERR_ARG, "null argument");
  *out = nullptr;
  if (sel < 0 || sel > 3) return set_error(GMBE_ERR_ARG, "sel must be 0..3 (got %d)", sel);
  if (ngpus != 1 || sel == 3)
    return set_error(GMBE_ERR_UNSUPPORTED,
                     "this engine runs on one GPU (sel=%d ngpus=%d)",
                     sel, ngpus);
  if (bound_height < 0 || bound_size < 0)
    return set_error(GMBE_ERR_ARG, "bounds must be non-negative");
  gmbe_finder* f = new gmbe_finder();
  f->g = g;
  f->sel = sel;
  f->bound_height = bound_height;
  f->bound_size = bound_size;
  int rc = f->eng.ensure_device(0);
  if (rc) {
    delete f;
    return rc;
  }
  if (hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&f->ev0) != hipSuccess || hipEventCreate(&f->ev1) != hipSuccess ||
      hipEventCreate(&f->evm) != hipSuccess) {
    delete f;
    return set_error(GMBE_ERR_HIP, "stream/event creation failed");
  }
  rc = f->dg.upload(*g);
  if (rc == GMBE_OK &&
      hipMalloc(&f->ctl, C_NSLOTS * kCtlStride * sizeof(unsigned long long)) != hipSuccess)
    rc = set_error(GMBE_ERR_NOMEM, "control block");
  if (rc == GMBE_OK && hipMalloc(&f->acc, mbe::A_COUNT * sizeof(unsigned long long)) != hipSuccess)
    rc = set_error(GMBE_ERR_NOMEM, "result block");
  if (rc == GMBE_OK && hipMalloc(&f->gt, sizeof(mbe::DevGraph)) != hipSuccess)
    rc = set_error(GMBE_ERR_NOMEM, "graph table");
  if (rc) {
    gmbe_finder_destroy(f);
    return rc;
  }
  *out = f;
  return GMBE_OK;
}

int gmbe_finder_create(const gmbe_graph* g, int sel, int bound_height, int bound_size, int ngpus,
                       gmbe_finder** out) {
  return gmbe::guarded("gmbe_finder_create", [&] {
    return gmbe_finder_create_impl(g, sel, bound_height, bound_size, ngpus, out);
  });
}

static int gmbe_finder_set_impl(gmbe_finder* f, const char* key, int64_t value) {
  if (!f || !key) return set_error(GMBE_ERR_ARG, "null argument");
  std::string k(key);
  auto range = [&](int64_t lo, int64_t hi) -> int {
    if (value < lo || value > hi)
      return set_error(GMBE_ERR_ARG, "%s=%lld out of range [%lld, %lld]", key, (long long)value,
                       (long long)lo, (long long)hi);
    return GMBE_OK;
  };
  int rc;
  if (k == "waves_per_block") {
    if (value != 2 && value != 4 && value != 8)
      return set_error(GMBE_ERR_ARG, "waves_per_block must be 2, 4 or 8");
    f->waves_per_block = value;
  }
  else if (k == "blocks") { if ((rc = range(0, 1 << 20))) return rc; f->blocks = value; }
  else if (k == "prune") { if ((rc = range(0, 1))) return rc; f->prune = value; }
  else if (k == "max_prefix") { if ((rc = range(2, mbe::kMaxPrefix))) return rc; f->max_prefix = value; }
  else if (k == "lds_kb") {
    if (value != 0 && value != 3 && value != 7 && value != 16)
      return set_error(GMBE_ERR_ARG, "lds_kb must be 0, 3, 7 or 16");
    f->lds_kb = value;
  }
  else if (k == "queue_log2") { if ((rc = range(0, 30))) return rc; f->queue_log2 = value; }
  else if (k == "list_cap") { if ((rc = range(0, (int64_t)1 << 34))) return rc; f->list_cap = value; }
  else if (k == "arena_kb") { if ((rc = range(64, 1 << 20))) return rc; f->arena_kb = value; f->arena_set = true; }
  else if (k == "root_stride") { if ((rc = range(1, 1 << 30))) return rc; f->root_stride = value; }
  else if (k == "share_work") { if ((rc = range(1, 1ll << 40))) return rc; f->share_work = value; }
  else if (k == "share_queue") { if ((rc = range(-1, 1ll << 40))) return rc; f->share_queue = value; }
  else if (k == "merge_branch") { if ((rc = range(0, 1ll << 30))) return rc; f->merge_branch = value; }
  else if (k == "merge_min") { if ((rc = range(2, 1ll << 30))) return rc; f->merge_min = value; }
  else if (k == "share") { if ((rc = range(0, 1))) return rc; f->share = value; }
  else if (k == "ctx_kb") { if ((rc = range(16, 1 << 20))) return rc; f->ctx_kb = value; f->ctx_set = true; }
  else if (k == "group") { if ((rc = range(0, 1))) return rc; f->group = value; }
  else if (k == "profile") { if ((rc = range(0, 1))) return rc; f->profile = value; }
  else if (k == "ctx_slots") { if ((rc = range(1, 32))) return rc; f->ctx_slots = value; }
  else if (k == "exit_stamps") { if ((rc = range(0, 1))) return rc; f->exit_stamps = value; }
  else if (k == "late_share") { if ((rc = range(0, 1ll << 30))) return rc; f->late_share = value; }
  else if (k == "dom_k") { if ((rc = range(0, mbe::kMaxDom))) return rc; f->dom_k = value; }
  else if (k == "dom_skip") { if ((rc = range(0, 1))) return rc; f->dom_skip = value; }
  else if (k == "recompute") { if ((rc = range(0, 1))) return rc; f->recompute = value; }
  else if (k == "count_bytes") { if ((rc = range(0, 1))) return rc; f->count_bytes = value; }
  else if (k == "reduce_min") { if ((rc = range(0, 1 << 20))) return rc; f->reduce_min = value; }
  else if (k == "wide_resident") { if ((rc = range(0, 1))) return rc; f->wide_resident = value; }
  else if (k == "deadline_ms") { if ((rc = range(0, 1ll << 32))) return rc; f->deadline_ms = value; }
  else if (k == "root_begin") { if ((rc = range(0, (1ll << 32)))) return rc; f->root_begin = value; }
  else if (k == "root_end") { if ((rc = range(-1, (1ll << 32)))) return rc; f->root_end = value; }
  else if (k == "root_phase") { if ((rc = range(0, (1 << 30) - 1))) return rc; f->root_phase = value; }
  else if (k == "bound_height") { if ((rc = range(0, 1 << 30))) return rc; f->bound_height = (int)value; }
  else if (k == "bound_size") { if ((rc = range(0, (int64_t)1 << 30))) return rc; f->bound_size = (int)value; }
  else return set_error(GMBE_ERR_ARG, "unknown knob '%s'", key);
  return GMBE_OK;
}

int gmbe_finder_set(gmbe_finder* f, const char* key, int64_t value) {
  return gmbe::guarded("gmbe_finder_set", [&] { return gmbe_finder_set_impl(f, key, value); });
}

static int gmbe_finder_execute_impl(gmbe_finder* f, gmbe_result* res) {
  if (!f || !res) return set_error(GMBE_ERR_ARG, "null argument");
  if (f->root_phase >= f->root_stride)
    return set_error(GMBE_ERR_ARG, "root_phase must be < root_stride");
  memset(res, 0, sizeof(*res));
  HIP_TRY(hipSetDevice(f->eng.device));
  int rc = finder_alloc(f);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(f->stream));
  double t0 = now_s();  // start_time_ (src/IterFinderGpu.cu:1096)
  // roots: processing ranks r in [root_begin, root_end) with r % stride == phase
  const uint64_t nV = f->dg.nV, stride = (uint64_t)f->root_stride;
  const uint64_t end = f->root_end < 0 ? nV : std::min<uint64_t>((uint64_t)f->root_end, nV);
  const uint64_t b = (uint64_t)f->root_begin;
  const uint64_t first = b + ((uint64_t)f->root_phase + stride - b % stride) % stride;
  const uint64_t n_roots = end > first ? (end - first + stride - 1) / stride : 0;
  std::vector<unsigned long long> ctl_h(C_NSLOTS * kCtlStride, 0ull);
  ctl_h[C_WORK_LEFT * kCtlStride] = n_roots;
  HIP_TRY(hipMemcpyAsync(f->ctl, ctl_h.data(), ctl_h.size() * 8, hipMemcpyHostToDevice,
                         f->stream));
  HIP_TRY(hipMemsetAsync(f->acc, 0, mbe::A_COUNT * 8, f->stream));
  hipLaunchKernelGGL(mbe::init_queue, dim3(256), dim3(256), 0, f->stream, f->qseq,
                     (unsigned long long)f->qcap);
  HIP_TRY(hipGetLastError());
  mbe::Params P{};
  {
    // the graph table of a finder: one entry
    mbe::DevGraph e = f->dg.view();
    e.bound_height = f->bound_height;
    e.bound_size = f->bound_size;
    HIP_TRY(hipMemcpyAsync(f->gt, &e, sizeof(e), hipMemcpyHostToDevice, f->stream));
  }
  P.gt = f->gt;
  P.per_graph = nullptr;
  P.s.root_next = f->ctl + C_ROOT_NEXT * kCtlStride;
  P.s.work_left = f->ctl + C_WORK_LEFT * kCtlStride;
  P.s.qctl = f->ctl + C_QCTL * kCtlStride;
  P.s.queued = f->ctl + C_QUEUED * kCtlStride;
  P.list_cursor = f->ctl + C_LIST_CURSOR * kCtlStride;
  P.s.n_deferred = f->ctl + C_DEFERRED * kCtlStride;
  P.s.deadline_word = f->ctl + C_DEADLINE * kCtlStride;
  P.s.deferred = f->deferred;
  P.s.deferred_cap = f->deferred_cap;
  P.s.q = f->q;
  P.s.qseq = f->qseq;
  P.s.qcap = f->qcap;
  P.s.n_roots = n_roots;
  P.s.stride = (uint32_t)stride;
  P.s.phase = (uint32_t)first;
  P.s.root_tasks = nullptr;
  P.w = f->layout;
  P.acc = f->acc;
  P.list = f->list;
  P.list_cap = f->list_words;
  P.prune = (int)f->prune;
  P.max_prefix = (int)f->max_prefix;
  P.split = f->sel == 2;
  P.profile = (int)f->profile;
  P.group = (int)f->group;
  P.share = f->share && f->sel == 2;  // sharing splits a task: only for -s 2
  P.merge_min = (int)f->merge_min;
  P.merge_branch = (int)f->merge_branch;
  P.budget = (uint64_t)f->deadline_ms * (uint64_t)f->eng.wall_khz;
  P.exit_stamps = nullptr;
  if (f->exit_stamps) {
    if (f->stamps) (void)hipFree(f->stamps);
    f->stamps = nullptr;
    HIP_TRY(hipMalloc(&f->stamps, 2ull * f->nblocks * sizeof(unsigned long long)));
    HIP_TRY(hipMemsetAsync(f->stamps, 0, 2ull * f->nblocks * sizeof(unsigned long long), f->stream));
    P.exit_stamps = f->stamps;
  }
  P.late_share = f->share && f->sel == 2 ? (int)f->late_share : 0;  // a form of splitting
  P.dom_k = (int)f->dom_k;
  P.dom_skip = (int)f->dom_skip;
  P.recompute = (int)f->recompute;
  P.count_bytes = (int)f->count_bytes;
  P.reduce_min = (int)f->reduce_min;
  P.share_work = (int)std::min<int64_t>(f->share_work, INT32_MAX);
  P.share_queue = f->share_queue >= 0 ? (int)f->share_queue : f->nwaves;
  KernelChoice kc = pick_kernel((int)f->lds_kb, (int)f->waves_per_block, false);
  hipLaunchKernelGGL(mbe::reset_ctx, dim3(64), dim3(256), 0, f->stream, P,
                     (unsigned long long)f->nwaves * (unsigned long long)f->ctx_slots);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(f->ev0, f->stream));
  if (n_roots > 0) {
    hipLaunchKernelGGL(kc.fn, dim3(f->nblocks), dim3(kc.waves * 64), 0, f->stream, P);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipEventRecord(f->evm, f->stream));
  std::vector<unsigned long long> ctl_out(C_NSLOTS * kCtlStride, 0ull);
  HIP_TRY(hipMemcpyAsync(ctl_out.data(), f->ctl, ctl_out.size() * 8, hipMemcpyDeviceToHost,
                         f->stream));
  HIP_TRY(hipMemcpyAsync(f->main_acc, f->acc, sizeof(f->main_acc), hipMemcpyDeviceToHost,
                         f->stream));
  HIP_TRY(hipStreamSynchronize(f->stream));
  f->last_deferred = std::min<uint64_t>(ctl_out[C_DEFERRED * kCtlStride], f->deferred_cap);
  // the wide pass gets what is left of the budget
  const double used_ms = (now_s() - t0) * 1e3;
  const bool out_of_time = f->deadline_ms > 0 && used_ms >= (double)f->deadline_ms;
  if (f->deadline_ms > 0 && !out_of_time)
    P.budget = (uint64_t)(((double)f->deadline_ms - used_ms) * f->eng.wall_khz);
  if (f->last_deferred > 0 && !out_of_time) {
    // wide-bitset pass over the deferred tasks (same queue, fresh counters)
    std::vector<unsigned long long> ctl2(C_NSLOTS * kCtlStride, 0ull);
    ctl2[C_WORK_LEFT * kCtlStride] = f->last_deferred;
    ctl2[C_LIST_CURSOR * kCtlStride] = ctl_out[C_LIST_CURSOR * kCtlStride];
    ctl2[C_DEADLINE * kCtlStride] = ctl_out[C_DEADLINE * kCtlStride];
    HIP_TRY(hipMemcpyAsync(f->ctl, ctl2.data(), ctl2.size() * 8, hipMemcpyHostToDevice,
                           f->stream));
    hipLaunchKernelGGL(mbe::init_queue, dim3(256), dim3(256), 0, f->stream, f->qseq,
                       (unsigned long long)f->qcap);
    P.s.root_tasks = f->deferred;
    P.s.n_roots = f->last_deferred;
    P.exit_stamps = nullptr;  // the stamps describe the main pass
    P.s.deferred = nullptr;
    P.s.deferred_cap = 0;
    KernelChoice wk = pick_kernel(8, 4, true);
    // the wide kernel holds fewer waves per SIMD than the main one: launch the
    // blocks that are resident at once (every wave index stays within the slabs)
    const int wblocks = f->wide_resident
                            ? std::min(f->nblocks, resident_blocks(wk, f->eng.cu_count))
                            : f->nblocks;
    hipLaunchKernelGGL(wk.fn, dim3(wblocks), dim3(wk.waves * 64), 0, f->stream, P);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipEventRecord(f->ev1, f->stream));
  unsigned long long acc_h[mbe::A_COUNT];
  HIP_TRY(hipMemcpyAsync(acc_h, f->acc, sizeof(acc_h), hipMemcpyDeviceToHost, f->stream));
  HIP_TRY(hipMemcpyAsync(ctl_out.data(), f->ctl, ctl_out.size() * 8, hipMemcpyDeviceToHost,
                         f->stream));
  HIP_TRY(hipStreamSynchronize(f->stream));
  double t1 = now_s();
  float ms = 0, mms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, f->ev0, f->ev1));
  HIP_TRY(hipEventElapsedTime(&mms, f->ev0, f->evm));
  f->main_ms = mms;
  f->wide_ms = f->last_deferred > 0 ? (double)ms - mms : 0.0;
  res->maximal = acc_h[mbe::A_MAXIMAL];
  res->non_maximal = acc_h[mbe::A_NONMAX];
  res->hash_sum = acc_h[mbe::A_HSUM];
  res->hash_xor = acc_h[mbe::A_HXOR];
  res->nodes = acc_h[mbe::A_NODES];
  res->tasks = acc_h[mbe::A_TASKS];
  res->splits = acc_h[mbe::A_SPLITS];
  res->overflow = acc_h[mbe::A_OVERFLOW];
  res->alg_bytes = acc_h[mbe::A_BALG];
  res->exe_time_s = t1 - t0;
  res->kernel_time_s = ms * 1e-3;
  f->ran = true;
  f->list_used = std::min<uint64_t>(ctl_out[C_LIST_CURSOR * kCtlStride], f->list_words);
  f->list_ovf = acc_h[mbe::A_LIST_OVF];
  memcpy(f->prof, acc_h, sizeof(acc_h));
  f->exit_s.clear();
  if (P.exit_stamps && n_roots > 0) {
    std::vector<unsigned long long> st(2ull * f->nblocks);
    HIP_TRY(hipMemcpy(st.data(), f->stamps, st.size() * 8, hipMemcpyDeviceToHost));
    unsigned long long t0s = ~0ull;
    for (int b = 0; b < f->nblocks; ++b) t0s = std::min<unsigned long long>(t0s, st[2 * b]);
    for (int b = 0; b < f->nblocks; ++b)
      f->exit_s.push_back((double)(st[2 * b + 1] - t0s) / ((double)f->eng.wall_khz * 1e3));
  }
  const uint64_t roots_claimed = std::min<uint64_t>(ctl_out[C_ROOT_NEXT * kCtlStride], n_roots);
  if (f->deadline_ms > 0 &&
      (acc_h[mbe::A_ABANDONED] > 0 || ctl_out[C_WORK_LEFT * kCtlStride] != 0 ||
       (f->last_deferred > 0 && out_of_time)))
    return set_error(GMBE_ERR_DEADLINE,
                     "deadline of %lld ms passed: %llu of %llu top-level vertices claimed, %llu "
                     "subtree(s) abandoned; the counts are partial",
                     (long long)f->deadline_ms, (unsigned long long)roots_claimed,
                     (unsigned long long)n_roots, (unsigned long long)acc_h[mbe::A_ABANDONED]);
  if (res->overflow)
    return set_error(GMBE_ERR_CAPACITY,
                     "%llu task(s) exceeded a capacity (%llu list arena (arena_kb=%lld), %llu |L| > "
                     "%d, %llu deferred-buffer); the result is incomplete",
                     (unsigned long long)res->overflow, (unsigned long long)acc_h[mbe::A_OVF_ARENA],
                     (long long)f->arena_kb, (unsigned long long)acc_h[mbe::A_OVF_WIDE],
                     mbe::kMaxW * 64, (unsigned long long)acc_h[mbe::A_OVF_DEFER]);
  return GMBE_OK;
}

int gmbe_finder_execute(gmbe_finder* f, gmbe_result* res) {
  return gmbe::guarded("gmbe_finder_execute", [&] { return gmbe_finder_execute_impl(f, res); });
}

static int gmbe_finder_exit_times_impl(const gmbe_finder* f, double* out, size_t cap, size_t* n) {
  if (!f || !n) return set_error(GMBE_ERR_ARG, "null argument");
  if (f->exit_s.empty()) return set_error(GMBE_ERR_STATE, "no run with exit_stamps set");
  *n = f->exit_s.size();
  if (out) memcpy(out, f->exit_s.data(), std::min(cap, f->exit_s.size()) * sizeof(double));
  return GMBE_OK;
}

int gmbe_finder_exit_times(const gmbe_finder* f, double* out, size_t cap, size_t* n) {
  return gmbe::guarded("gmbe_finder_exit_times", [&] {
    return gmbe_finder_exit_times_impl(f, out, cap, n);
  });
}

static int gmbe_finder_list_impl(gmbe_finder* f, uint32_t* buf, size_t cap, size_t* used) {
  if (!f || !used) return set_error(GMBE_ERR_ARG, "null argument");
  if (!f->ran || !f->list) return set_error(GMBE_ERR_STATE, "no list-mode run (set list_cap)");
  if (f->list_ovf)
    return set_error(GMBE_ERR_CAPACITY, "device list buffer overflowed (list_cap=%lld)",
                     (long long)f->list_cap);
  std::vector<uint32_t> raw(f->list_used);
  if (f->list_used)
    HIP_TRY(hipMemcpy(raw.data(), f->list, f->list_used * 4, hipMemcpyDeviceToHost));
  *used = f->list_used;
  if (!buf) return GMBE_OK;
  if (cap < f->list_used) return set_error(GMBE_ERR_CAPACITY, "buffer too small");
  // internal (L, R) -> original (A, B)
  const gmbe_graph* g = f->g;
  size_t i = 0, o = 0;
  while (i < raw.size()) {
    uint32_t nL = raw[i], nR = raw[i + 1];
    const uint32_t* L = &raw[i + 2];
    const uint32_t* R = L + nL;
    if (!g->swapped) {
      buf[o++] = nL;
      buf[o++] = nR;
      for (uint32_t j = 0; j < nL; ++j) buf[o++] = g->u_orig[L[j]];
      for (uint32_t j = 0; j < nR; ++j) buf[o++] = g->v_orig[R[j]];
    } else {
      buf[o++] = nR;
      buf[o++] = nL;
      for (uint32_t j = 0; j < nR; ++j) buf[o++] = g->v_orig[R[j]];
      for (uint32_t j = 0; j < nL; ++j) buf[o++] = g->u_orig[L[j]];
    }
    i += 2 + (size_t)nL + nR;
  }
  return GMBE_OK;
}

int gmbe_finder_list(gmbe_finder* f, uint32_t* buf, size_t cap, size_t* used) {
  return gmbe::guarded("gmbe_finder_list", [&] {
    return gmbe_finder_list_impl(f, buf, cap, used);
  });
}

static int gmbe_finder_stats_impl(gmbe_finder* f, uint64_t* out, size_t n) {
  if (!f) return set_error(GMBE_ERR_ARG, "null finder");
  // the profile counters, then the main kernel's and the wide pass's time (ns),
  // the tasks handed to the wide pass, and the nodes, maximal bicliques,
  // algorithmic bytes and list bytes of the main kernel alone
  const int first = mbe::A_CYC_WAIT, nprof = mbe::A_COUNT - mbe::A_CYC_WAIT;
  uint64_t all[mbe::A_COUNT + 7];
  for (int i = 0; i < nprof; ++i) all[i] = f->prof[first + i];
  all[nprof] = (uint64_t)(f->main_ms * 1e6);
  all[nprof + 1] = (uint64_t)(f->wide_ms * 1e6);
  all[nprof + 2] = f->last_deferred;
  all[nprof + 3] = f->main_acc[mbe::A_NODES];
  all[nprof + 4] = f->main_acc[mbe::A_MAXIMAL];
  all[nprof + 5] = f->main_acc[mbe::A_BALG];
  all[nprof + 6] = f->main_acc[mbe::A_BLIST];
  const int avail = nprof + 7;
  for (size_t i = 0; out && i < n && i < (size_t)avail; ++i) out[i] = all[i];
  return avail;
}

int gmbe_finder_stats(gmbe_finder* f, uint64_t* out, size_t n) {
  return gmbe::guarded("gmbe_finder_stats", [&] { return gmbe_finder_stats_impl(f, out, n); });
}

void gmbe_finder_destroy(gmbe_finder* f) {
  if (!f) return;
  if (f->eng.device >= 0) (void)hipSetDevice(f->eng.device);
  if (f->stream) (void)hipStreamSynchronize(f->stream);
  f->dg.release();
  (void)hipFree(f->scratch);
  (void)hipFree(f->q);
  (void)hipFree(f->qseq);
  (void)hipFree(f->ctl);
  (void)hipFree(f->acc);
  (void)hipFree(f->gt);
  (void)hipFree(f->stamps);
  (void)hipFree(f->list);
  (void)hipFree(f->deferred);
  if (f->ev0) (void)hipEventDestroy(f->ev0);
  if (f->ev1) (void)hipEventDestroy(f->ev1);
  if (f->evm) (void)hipEventDestroy(f->evm);
  if (f->stream) (void)hipStreamDestroy(f->stream);
  delete f;
}

static int gmbe_device_info_impl(char* buf, size_t cap, int* cu_count) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0)
    return set_error(GMBE_ERR_HIP, "no HIP device available (%s)", hipGetErrorString(e));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, 0));
  if (buf && cap) snprintf(buf, cap, "%s (%s)", prop.name, prop.gcnArchName);
  if (cu_count) *cu_count = prop.multiProcessorCount;
  return GMBE_OK;
}

int gmbe_device_info(char* buf, size_t cap, int* cu_count) {
  return gmbe::guarded("gmbe_device_info", [&] {
    return gmbe_device_info_impl(buf, cap, cu_count);
  });
}

}  // extern "C"
