// device_graph.hpp — device-resident CSR shared by the finder, batch and
// serving forms (CSRBiGraph::CopyToGpu / Reset, src/BiGraph.cu:413-454).
#pragma once

#include <hip/hip_runtime.h>

#include "graph.hpp"
#include "mbe_kernel.hpp"

namespace gmbe {

struct Engine {
  int device = -1;
  int cu_count = 0;
  int wall_khz = 100000;  // rate of the device's constant wall clock
  int ensure_device(int dev);
};

struct DeviceGraph {
  uint32_t nU = 0, nV = 0, dmaxV = 0;
  uint64_t E = 0;
  int swapped = 0;
  int64_t *v_off = nullptr, *u_off = nullptr;
  uint32_t *v_adj = nullptr, *u_adj = nullptr;
  uint64_t *kUd = nullptr, *kVd = nullptr;
  int upload(const gmbe_graph& g);
  void release();
  mbe::DevGraph view() const;
};

mbe::Scratch scratch_layout(uint32_t nV, uint32_t dmaxV, uint64_t arena_bytes,
                            uint64_t ctx_bytes, uint32_t ctx_slots);

// Control counters live kCtlStride u64 (256 bytes) apart so that polling one
// never contends with atomics on another; the shards' head/tail pairs follow.
enum CtlSlot { C_ROOT_NEXT = 0, C_WORK_LEFT, C_QUEUED, C_LIST_CURSOR, C_DEFERRED, C_WIDE_N,
               C_DEADLINE,  // absolute deadline of the run (wall clock), shared by both passes
               C_QCTL = 8, C_NSLOTS = 8 + 2 * mbe::kShards };
constexpr uint64_t kCtlWords = (uint64_t)C_NSLOTS * mbe::kCtlStride;

typedef void (*KernelFn)(mbe::Params);
struct KernelChoice {
  KernelFn fn;
  int waves;
  int lds_per_wave;
};

// Enumeration kernel variant: waves per block (2, 4 or 8) and LDS list arena
// per wave in KB; `wide` selects the 16-word-bitset kernel of the second pass.
// Returns {nullptr} for a combination that is not built.
KernelChoice pick_kernel(int lds_kb, int waves, bool wide);
// the variants (kern_*.hip)
KernelChoice kernel_w4_l7();
KernelChoice kernel_w4_l3();
KernelChoice kernel_w4_l0();
KernelChoice kernel_w4_l16();
KernelChoice kernel_w8_l7();
KernelChoice kernel_w2_l7();
KernelChoice kernel_wide();
// Resident blocks of a variant on this device (persistent grid size).
int resident_blocks(const KernelChoice& kc, int cu_count);

// Stream-ordered launches (recordable into a HIP graph: no memset nodes).
hipError_t launch_persistent(const KernelChoice& kc, int blocks, hipStream_t st,
                             const mbe::Params& P);
hipError_t launch_init_queue(hipStream_t st, unsigned long long* seq, unsigned long long cap);
hipError_t launch_reset_ctx(hipStream_t st, const mbe::Params& P, unsigned long long n);
// Per-graph outstanding units of a batch = each graph's top-level vertices.
hipError_t launch_init_left(hipStream_t st, const mbe::Params& P, unsigned long long n);
hipError_t launch_fill(hipStream_t st, unsigned long long* p, unsigned long long n,
                       unsigned long long v);
// Control block for a new launch: zero, then work_left = n_roots; with
// keep_deadline the run's deadline word stays (a later launch of one run).
hipError_t launch_ctl_start(hipStream_t st, unsigned long long* ctl, unsigned long long n_roots,
                            bool keep_deadline = false);
// Control block for the wide pass over the tasks the main pass deferred:
// C_WIDE_N = min(n_deferred, cap), work_left = C_WIDE_N, the list cursor and
// the deadline kept, the rest zero.
hipError_t launch_ctl_wide(hipStream_t st, unsigned long long* ctl, unsigned long long cap);

}  // namespace gmbe
