// static_entry.hip — the one entry point that allocates nothing
// (gmbe_finder_execute_static).  It has no counterpart in the reference: it
// serves callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).  The graph
// copy, the control block, the queue and the per-wave slabs of one
// enumeration are carved out of a file-scope device array of kStaticBytes;
// the kernels are the finder's, on a grid sized to what fits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>

#include "device_graph.hpp"
#include "graph.hpp"
#include "mbe_kernel.hpp"

using namespace gmbe;

namespace {

constexpr uint64_t kStaticBytes = 128ull << 20;  // 128 MiB per GPU, held until exit
constexpr int kMaxDevices = 64;
constexpr uint32_t kStaticWaves = 256;     // at most one block of four waves per CU
constexpr uint64_t kStaticQcap = 64;       // queue slots per shard
constexpr uint64_t kStaticDeferred = 4096; // tasks the wide pass can take
constexpr uint64_t kStaticArena = 512ull << 10;
constexpr uint64_t kStaticCtx = 128ull << 10;
constexpr uint32_t kStaticCtxSlots = 2;

static __device__ __align__(256) unsigned char scratch[kStaticBytes];

// the array's device address on each device (taken once per device)
unsigned char* g_ptr[kMaxDevices] = {};
std::mutex g_mu;

uint64_t up(uint64_t x) { return (x + 255) & ~255ull; }

}  // namespace

namespace mbe {
// Hash keys of one side from its original ids (the host's hash_keys, on the
// device, so that nothing is staged in host memory).
__global__ void static_keys(uint64_t* k, const uint32_t* orig, uint32_t n, uint32_t side) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    // the splitmix64 finalizer of gmbe::side_key
    uint64_t z = (((uint64_t)orig[i] << 1) | side) + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    k[i] = z ^ (z >> 31);
  }
}
// The graph table entry of the static run.
__global__ void static_table(DevGraph* gt, DevGraph e) { *gt = e; }
}  // namespace mbe

#define S_TRY(expr)                                                                           \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return set_error(GMBE_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),  \
                       __FILE__, __LINE__);                                                   \
  } while (0)

// The device address of `scratch` on the current device.
static int static_ptr(unsigned char** out) {
  int dev = 0;
  S_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return set_error(GMBE_ERR_UNSUPPORTED, "device %d", dev);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    S_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_ptr[dev] = (unsigned char*)p;
  }
  *out = g_ptr[dev];
  return GMBE_OK;
}

extern "C" {

static int execute_static_impl(const gmbe_graph* g, int bound_height, int bound_size,
                               gmbe_result* res) {
  if (!g || !res) return set_error(GMBE_ERR_ARG, "null argument");
  if (bound_height < 0 || bound_size < 0)
    return set_error(GMBE_ERR_ARG, "bounds must be non-negative");
  memset(res, 0, sizeof(*res));
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return set_error(GMBE_ERR_HIP, "no HIP device available");
  unsigned char* base = nullptr;
  int rc = static_ptr(&base);
  if (rc) return rc;
  int dev = 0;
  S_TRY(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  S_TRY(hipGetDeviceProperties(&prop, dev));
  const uint64_t nU = g->nU, nV = g->nV, E = g->E;
  uint32_t dmaxV = 0;
  for (uint64_t v = 0; v < nV; ++v)
    dmaxV = std::max<uint32_t>(dmaxV, (uint32_t)(g->v_off[v + 1] - g->v_off[v]));
  // carve the array: graph, its keys and ids, table, control, results, queue,
  // deferred tasks, then as many per-wave slabs as fit
  uint64_t o = 0;
  auto take = [&](uint64_t bytes) {
    const uint64_t r = o;
    o = up(o + std::max<uint64_t>(bytes, 1));
    return r;
  };
  const uint64_t o_voff = take((nV + 1) * 8), o_uoff = take((nU + 1) * 8);
  const uint64_t o_vadj = take(E * 4), o_uadj = take(E * 4);
  const uint64_t o_kU = take(nU * 8), o_kV = take(nV * 8);
  const uint64_t o_uorig = take(nU * 4), o_vorig = take(nV * 4);
  const uint64_t o_gt = take(sizeof(mbe::DevGraph));
  const uint64_t o_ctl = take(kCtlWords * 8), o_acc = take(mbe::A_COUNT * 8);
  const uint64_t o_q = take(kStaticQcap * mbe::kShards * mbe::kEntryWords * 4);
  const uint64_t o_qseq = take(kStaticQcap * mbe::kShards * 8);
  const uint64_t o_def = take(kStaticDeferred * sizeof(mbe::Task));
  mbe::Scratch lay = scratch_layout((uint32_t)nV, dmaxV, kStaticArena, kStaticCtx,
                                    kStaticCtxSlots);
  const uint64_t o_slabs = up(o);
  const uint64_t room = o_slabs < kStaticBytes ? kStaticBytes - o_slabs : 0;
  const KernelChoice kc = pick_kernel(7, 4, false), wk = pick_kernel(8, 4, true);
  if (!kc.fn || !wk.fn || kc.waves != wk.waves)
    return set_error(GMBE_ERR_INTERNAL, "kernel variants of the static run are not built");
  const uint64_t resident = (uint64_t)resident_blocks(kc, prop.multiProcessorCount) * kc.waves;
  uint64_t nwaves = std::min<uint64_t>(kStaticWaves, resident);
  nwaves = std::min<uint64_t>(nwaves, room / lay.slab) / kc.waves * kc.waves;
  if (nwaves < (uint64_t)kc.waves)
    return set_error(GMBE_ERR_CAPACITY,
                     "graph too large for the static scratch: |U|=%llu |V|=%llu |E|=%llu need "
                     "%llu bytes for the graph and one block of slabs of %llu; the array holds %llu",
                     (unsigned long long)nU, (unsigned long long)nV, (unsigned long long)E,
                     (unsigned long long)o_slabs, (unsigned long long)lay.slab,
                     (unsigned long long)kStaticBytes);
  const uint32_t nblocks = (uint32_t)(nwaves / kc.waves);
  lay.base = (char*)(base + o_slabs);
  hipStream_t st = nullptr;  // the null stream: no stream is created
  const double t0 = now_s();
  // the graph (host -> the array), its keys computed on the device
  S_TRY(hipMemcpy(base + o_voff, g->v_off.data(), (nV + 1) * 8, hipMemcpyHostToDevice));
  S_TRY(hipMemcpy(base + o_uoff, g->u_off.data(), (nU + 1) * 8, hipMemcpyHostToDevice));
  if (E) {
    S_TRY(hipMemcpy(base + o_vadj, g->v_adj.data(), E * 4, hipMemcpyHostToDevice));
    S_TRY(hipMemcpy(base + o_uadj, g->u_adj.data(), E * 4, hipMemcpyHostToDevice));
  }
  if (nU) S_TRY(hipMemcpy(base + o_uorig, g->u_orig.data(), nU * 4, hipMemcpyHostToDevice));
  if (nV) S_TRY(hipMemcpy(base + o_vorig, g->v_orig.data(), nV * 4, hipMemcpyHostToDevice));
  const uint32_t sideU = g->swapped ? 1 : 0, sideV = g->swapped ? 0 : 1;
  if (nU)
    hipLaunchKernelGGL(mbe::static_keys, dim3(64), dim3(256), 0, st, (uint64_t*)(base + o_kU),
                       (const uint32_t*)(base + o_uorig), (uint32_t)nU, sideU);
  if (nV)
    hipLaunchKernelGGL(mbe::static_keys, dim3(64), dim3(256), 0, st, (uint64_t*)(base + o_kV),
                       (const uint32_t*)(base + o_vorig), (uint32_t)nV, sideV);
  S_TRY(hipGetLastError());
  mbe::DevGraph e{};
  e.v_off = (const int64_t*)(base + o_voff);
  e.v_adj = (const uint32_t*)(base + o_vadj);
  e.u_off = (const int64_t*)(base + o_uoff);
  e.u_adj = (const uint32_t*)(base + o_uadj);
  e.kU = (const uint64_t*)(base + o_kU);
  e.kV = (const uint64_t*)(base + o_kV);
  e.nU = (uint32_t)nU;
  e.nV = (uint32_t)nV;
  e.dmaxV = dmaxV;
  e.swapped = (uint32_t)g->swapped;
  e.bound_height = bound_height;
  e.bound_size = bound_size;
  mbe::DevGraph* gt = (mbe::DevGraph*)(base + o_gt);
  hipLaunchKernelGGL(mbe::static_table, dim3(1), dim3(1), 0, st, gt, e);
  S_TRY(hipGetLastError());
  // a clean start: the slabs' count maps and bitmaps must be zero, and so
  // must every counter; fill kernels only (no memset)
  unsigned long long* ctl = (unsigned long long*)(base + o_ctl);
  unsigned long long* acc = (unsigned long long*)(base + o_acc);
  unsigned long long* qseq = (unsigned long long*)(base + o_qseq);
  S_TRY(launch_fill(st, (unsigned long long*)lay.base, nwaves * lay.slab / 8, 0ull));
  S_TRY(launch_fill(st, acc, mbe::A_COUNT, 0ull));
  S_TRY(launch_ctl_start(st, ctl, nV));
  S_TRY(launch_init_queue(st, qseq, kStaticQcap));
  mbe::Params P{};
  P.gt = gt;
  P.per_graph = nullptr;
  P.s.root_next = ctl + C_ROOT_NEXT * mbe::kCtlStride;
  P.s.work_left = ctl + C_WORK_LEFT * mbe::kCtlStride;
  P.s.qctl = ctl + C_QCTL * mbe::kCtlStride;
  P.s.queued = ctl + C_QUEUED * mbe::kCtlStride;
  P.list_cursor = ctl + C_LIST_CURSOR * mbe::kCtlStride;
  P.s.n_deferred = ctl + C_DEFERRED * mbe::kCtlStride;
  P.s.deferred = (mbe::Task*)(base + o_def);
  P.s.deferred_cap = kStaticDeferred;
  P.s.q = (uint32_t*)(base + o_q);
  P.s.qseq = qseq;
  P.s.qcap = kStaticQcap;
  P.s.n_roots = nV;
  P.s.stride = 1;
  P.s.phase = 0;
  P.s.root_tasks = nullptr;
  P.w = lay;
  P.acc = acc;
  P.list = nullptr;
  P.list_cap = 0;
  P.prune = 1;
  P.max_prefix = 4;
  P.split = 1;
  P.profile = 0;
  P.group = 1;
  P.share = 1;
  P.merge_min = 256;
  P.merge_branch = 16;
  P.budget = 0;
  P.exit_stamps = nullptr;
  P.late_share = 32;
  P.share_work = 1024;
  P.share_queue = (int)nwaves;
  P.dom_k = 8;
  P.dom_skip = 1;
  P.recompute = 0;
  P.count_bytes = 0;
  S_TRY(launch_reset_ctx(st, P, nwaves * kStaticCtxSlots));
  if (nV > 0) S_TRY(launch_persistent(kc, (int)nblocks, st, P));
  unsigned long long ctl_h[kCtlWords];
  S_TRY(hipMemcpy(ctl_h, ctl, sizeof(ctl_h), hipMemcpyDeviceToHost));
  const uint64_t nd = std::min<uint64_t>(ctl_h[C_DEFERRED * mbe::kCtlStride], kStaticDeferred);
  if (nd > 0) {
    // wide-bitset pass over the tasks the main pass deferred
    S_TRY(launch_ctl_wide(st, ctl, kStaticDeferred));
    S_TRY(launch_init_queue(st, qseq, kStaticQcap));
    P.s.root_tasks = P.s.deferred;
    P.s.n_roots = nd;
    P.s.deferred = nullptr;
    P.s.deferred_cap = 0;
    S_TRY(launch_persistent(wk, (int)nblocks, st, P));
  }
  unsigned long long acc_h[mbe::A_COUNT];
  S_TRY(hipMemcpy(acc_h, acc, sizeof(acc_h), hipMemcpyDeviceToHost));
  res->maximal = acc_h[mbe::A_MAXIMAL];
  res->non_maximal = acc_h[mbe::A_NONMAX];
  res->hash_sum = acc_h[mbe::A_HSUM];
  res->hash_xor = acc_h[mbe::A_HXOR];
  res->nodes = acc_h[mbe::A_NODES];
  res->tasks = acc_h[mbe::A_TASKS];
  res->splits = acc_h[mbe::A_SPLITS];
  res->overflow = acc_h[mbe::A_OVERFLOW];
  res->alg_bytes = 0;
  res->exe_time_s = now_s() - t0;
  res->kernel_time_s = res->exe_time_s;
  if (res->overflow)
    return set_error(GMBE_ERR_CAPACITY,
                     "%llu task(s) exceeded a capacity of the static run; the result is "
                     "incomplete", (unsigned long long)res->overflow);
  return GMBE_OK;
}

int gmbe_finder_execute_static(const gmbe_graph* g, int bound_height, int bound_size,
                               gmbe_result* res) {
  return gmbe::guarded("gmbe_finder_execute_static", [&] {
    return execute_static_impl(g, bound_height, bound_size, res);
  });
}

}  // extern "C"
