// persistent.hpp — the persistent enumeration kernel (template).  Each
// variant (waves per block, LDS list arena per wave, widest bitset, waves per
// SIMD) is instantiated in its own kern_*.hip so that they compile in
// parallel; pick_kernel (finder.hip) returns the one asked for.
#pragma once

#include "device_graph.hpp"
#include "mbe_kernel.hpp"

namespace mbe {

// Persistent grid: every wave loops over tasks until none are left.
// OCC: waves per SIMD the register allocation must leave room for (also
// imposed on the device functions the kernel calls).
template <int WAVES, int LDS_PER_WAVE, int WMAX, int OCC>
__global__ __launch_bounds__(WAVES * 64) __attribute__((amdgpu_waves_per_eu(OCC, 8)))
void mbe_persistent(Params P) {
  const uint64_t t_start = P.exit_stamps ? wall_clock64() : 0;
  __shared__ __attribute__((aligned(16))) char lds[WAVES * (LDS_PER_WAVE > 0 ? LDS_PER_WAVE : 16)];
  __shared__ uint32_t s_off[WAVES][kWave + 1];
  __shared__ int64_t s_start[WAVES][kWave];
  __shared__ Frame s_fr[WAVES][kLdsFrames];
  const uint32_t wv = threadIdx.x / kWave;
  const uint64_t gw = (uint64_t)blockIdx.x * WAVES + wv;
  WaveMem m = wave_mem(P, gw, lds + wv * (LDS_PER_WAVE > 0 ? LDS_PER_WAVE : 16), LDS_PER_WAVE,
                       s_off[wv], s_start[wv], s_fr[wv]);
  // the deadline: budget ticks after the first wave of the run started; with a
  // deadline word every later wave and the wide pass read the same absolute
  // time instead of arming their own
  uint64_t dl = 0;
  if (P.budget != 0) {
    dl = wall_clock64() + P.budget;
    if (P.s.deadline_word) {
      unsigned long long cur = 0;
      if (lane_id() == 0 &&
          !__hip_atomic_compare_exchange_strong(P.s.deadline_word, &cur, (unsigned long long)dl,
                                                __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT))
        dl = cur;
      dl = uni64(dl);
    }
  }
  m.deadline = dl;
  __shared__ Counters s_ctr[WAVES];
  Counters& c = s_ctr[wv];
  for (int i = (int)lane_id(); i < kCounters; i += kWave) as_lds(&c.maximal)[i].v = 0;
  wave_mem_wait();
  Task t;
  int k = 0;
  uint32_t word = 0;
  for (;;) {
    uint64_t tw = P.profile ? cycles() : 0;
    int kind = next_work(P, word, t, k, (uint32_t)(gw & (kShards - 1)), m.deadline);
    if (kind == 0) {
      // the final wait (nothing left for this wave) goes to the split slot's
      // neighbour so that mid-run waiting and the tail can be told apart
      if (P.profile) c.tail += cycles() - tw;
      break;
    }
    phase(c, P.profile, 0, tw);
    uint32_t gid;
    if (kind == 2) {
      gid = t.gid;
      m.g = P.gt + gid;
      Prefix pre{nullptr, 0, 0};
      if (k > 1) pre = prefix_common(*m.g, m, t, k - 1);
      run_task<WMAX>(P, m, t, k, pre, c, gw);
    } else if (read_lane(word, 0) & kItemKind) {
      // items of a shared task context (its header names the graph)
      gid = uni(ld_dev(&ctx_hdr(P, read_lane(word, 1))->gid));
      m.g = P.gt + gid;
      run_items<WMAX>(P, m, word, c, gw);
    } else {
      // an entry: children prefix + [c_i] of one split task; the prefix's
      // common neighbourhood is computed once for all of them
      const uint32_t hdr = read_lane(word, 0);
      const int kp = (int)child_hdr_k(hdr);
      const uint32_t cnt = child_hdr_cnt(hdr);
      gid = child_hdr_gid(hdr);
      m.g = P.gt + gid;
      for (int i = 0; i < kMaxPrefix; ++i) t.v[i] = i < kp ? read_lane(word, 1 + i) : kNone;
      t.gid = gid;
      const Prefix pre = prefix_common(*m.g, m, t, kp);
      for (uint32_t i = 0; i < cnt; ++i) {
        t.v[kp] = read_lane(word, 8 + i);
        run_task<WMAX>(P, m, t, kp + 1, pre, c, gw);
      }
    }
    if (P.per_graph) flush_graph(P, c, gid);
    work_done(P);
    graph_work(P, gid, -1);
  }
  if (P.exit_stamps) {
    // the reference's per-SM exit times (src/IterFinderGpu.cu:1038-1041): once
    // every wave of the block is done, the block's start and exit stamps
    __syncthreads();
    if (threadIdx.x == 0) {
      P.exit_stamps[2 * blockIdx.x] = t_start;
      P.exit_stamps[2 * blockIdx.x + 1] = wall_clock64();
    }
  }
  // the wave's counters into the result block (hash xor by xor, the rest by sum)
  wave_mem_wait();
  struct Slot { int acc; const LdsCount* v; };
  const Slot slots[] = {
      {A_MAXIMAL, &c.maximal}, {A_NONMAX, &c.nonmax}, {A_HSUM, &c.hsum}, {A_NODES, &c.nodes},
      {A_TASKS, &c.tasks}, {A_SPLITS, &c.splits}, {A_OVERFLOW, &c.overflow}, {A_BALG, &c.balg},
      {A_CYC_WAIT, &c.cyc[0]}, {A_CYC_PREFIX, &c.cyc[1]}, {A_CYC_EXPAND, &c.cyc[2]},
      {A_CYC_EXTRACT, &c.cyc[3]}, {A_CYC_BUILD, &c.cyc[4]}, {A_CYC_DFS, &c.cyc[5]},
      {A_CYC_SPLIT, &c.cyc[6]}, {A_EDGES, &c.edges}, {A_ENTRIES, &c.entries},
      {A_OVF_ARENA, &c.ovf_arena}, {A_OVF_WIDE, &c.ovf_wide}, {A_OVF_DEFER, &c.ovf_defer},
      {A_ENT_W1, &c.ent_w1}, {A_NODES_W1, &c.nodes_w1}, {A_CYC_TAIL, &c.tail},
      {A_CYC_PASS, &c.c_pass}, {A_CYC_MERGE, &c.c_merge}, {A_CYC_COMPRESS, &c.c_compress},
      {A_CYC_PUBLISH, &c.c_publish}, {A_PUBLISHES, &c.publishes}, {A_MERGES, &c.merges},
      {A_ABANDONED, &c.abandoned}, {A_LATE, &c.late}, {A_CYC_REDUCE, &c.c_reduce},
      {A_DROPPED, &c.dropped}, {A_SKIPPED, &c.skipped}, {A_BLIST, &c.blist}};
  for (const Slot& sl : slots) {
    const uint64_t v = sl.v->get();
    if (lane_id() == 0 && v != 0)
      __hip_atomic_fetch_add(P.acc + sl.acc, (unsigned long long)v, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
  }
  const uint64_t hx = c.hxor.get();
  if (lane_id() == 0 && hx != 0)
    __hip_atomic_fetch_xor(P.acc + A_HXOR, (unsigned long long)hx, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace mbe
