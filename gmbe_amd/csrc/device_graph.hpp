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

}  // namespace gmbe
