// wave.hpp — wave64 primitives for gfx950 (CDNA4).
//
// Replaces the warp32 helpers of src/GpuUtil.cuh:26-46 (lane id, popc, bfind
// in PTX) and :311-333 (shuffle broadcast, warp_sum/warp_min): here masks are
// 64-bit ballots, ordered compaction uses mbcnt, and uniform values are moved
// to scalar registers with readfirstlane.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mbe {

constexpr int kWave = 64;

__device__ __forceinline__ uint32_t lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}
__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }
// number of set bits of m in lanes below this one
__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
__device__ __forceinline__ uint32_t uni(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ int32_t uni(int32_t x) {
  return (int32_t)__builtin_amdgcn_readfirstlane((uint32_t)x);
}
__device__ __forceinline__ uint64_t uni64(uint64_t x) {
  uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
  uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint32_t read_lane(uint32_t x, uint32_t l) {
  return __builtin_amdgcn_readlane(x, l);
}
__device__ __forceinline__ uint64_t read_lane64(uint64_t x, uint32_t l) {
  uint32_t lo = __builtin_amdgcn_readlane((uint32_t)x, l);
  uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(x >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint32_t shfl_xor32(uint32_t x, int m) {
  return (uint32_t)__shfl_xor((int)x, m, kWave);
}
__device__ __forceinline__ uint64_t wave_sum64(uint64_t x) {
  for (int m = 32; m > 0; m >>= 1) {
    uint32_t lo = shfl_xor32((uint32_t)x, m), hi = shfl_xor32((uint32_t)(x >> 32), m);
    x += ((uint64_t)hi << 32) | lo;
  }
  return x;
}
__device__ __forceinline__ uint64_t wave_xor64(uint64_t x) {
  for (int m = 32; m > 0; m >>= 1) {
    uint32_t lo = shfl_xor32((uint32_t)x, m), hi = shfl_xor32((uint32_t)(x >> 32), m);
    x ^= ((uint64_t)hi << 32) | lo;
  }
  return x;
}
__device__ __forceinline__ uint32_t wave_sum32(uint32_t x) {
  for (int m = 32; m > 0; m >>= 1) x += shfl_xor32(x, m);
  return x;
}
// inclusive prefix sum across the wave
__device__ __forceinline__ uint32_t wave_incl_scan32(uint32_t x) {
  uint32_t l = lane_id();
  for (int d = 1; d < kWave; d <<= 1) {
    uint32_t y = (uint32_t)__shfl_up((int)x, d, kWave);
    if (l >= (uint32_t)d) x += y;
  }
  return x;
}
__device__ __forceinline__ int first_lane(uint64_t m) { return __builtin_ctzll(m); }
// This lane's bit of a uniform lane mask (32-bit shifts only).
__device__ __forceinline__ bool lane_bit(uint64_t m) {
  const uint32_t lane = lane_id();
  const uint32_t w = lane < 32u ? (uint32_t)m : (uint32_t)(m >> 32);
  return (w >> (lane & 31u)) & 1u;
}

// Maximum over the wave by DPP: row_shr 1/2/4/8 leave each row's maximum in
// its last lane, row_bcast:15 and row_bcast:31 carry rows into lane 63.
// Lanes whose DPP source lies outside the row read 0 (the identity).
__device__ __forceinline__ uint32_t dpp_max_step(uint32_t x, uint32_t y) { return x > y ? x : y; }
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t x) {
  x = dpp_max_step(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false));
  x = dpp_max_step(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false));
  x = dpp_max_step(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false));
  x = dpp_max_step(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false));
  x = dpp_max_step(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false));
  x = dpp_max_step(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false));
  return __builtin_amdgcn_readlane(x, 63);
}
// Minimum over the wave (as the maximum of the complement).
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t x) { return ~wave_max_u32(~x); }

// Minimum of (hi << 32 | lo) over the wave when no two lanes share the
// smallest hi (or all hold ~0): a 32-bit minimum, then the winner's lo.
__device__ __forceinline__ uint64_t wave_min_pair(uint64_t c) {
  const uint32_t hi = (uint32_t)(c >> 32);
  const uint32_t m = wave_min_u32(hi);
  const uint64_t at = __builtin_amdgcn_ballot_w64(hi == m);
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)c, (uint32_t)__builtin_ctzll(at));
  return ((uint64_t)m << 32) | lo;
}

// Sum over the wave by DPP, the same row_shr / row_bcast steps as the
// maximum, with 64-bit adds of the shifted halves.
template <int CTRL, int ROWS>
__device__ __forceinline__ uint64_t dpp_add64(uint64_t x) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)x, CTRL, ROWS, 0xf, false);
  const uint32_t hi =
      (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(x >> 32), CTRL, ROWS, 0xf, false);
  return x + (((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t x) {
  x = dpp_add64<0x111, 0xf>(x);
  x = dpp_add64<0x112, 0xf>(x);
  x = dpp_add64<0x114, 0xf>(x);
  x = dpp_add64<0x118, 0xf>(x);
  x = dpp_add64<0x142, 0xa>(x);
  x = dpp_add64<0x143, 0xc>(x);
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)x, 63);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(x >> 32), 63);
  return ((uint64_t)hi << 32) | lo;
}

// A generic pointer known to address LDS, as an LDS pointer: accesses through
// it are ds_ instructions instead of flat ones.
template <class T>
__device__ __forceinline__ __attribute__((address_space(3))) T* as_lds(T* p) {
  return (__attribute__((address_space(3))) T*)p;
}
template <class T>
__device__ __forceinline__ const __attribute__((address_space(3))) T* as_lds(const T* p) {
  return (const __attribute__((address_space(3))) T*)p;
}

// Ordering of this wave's own memory operations across its lanes.  Within a
// work-group (one CU) the hardware keeps them coherent, so this is a
// compiler-level fence (no cache maintenance is emitted).
__device__ __forceinline__ void wave_mem_wait() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
}
// Wait until every memory operation this wave issued has completed.  Used
// between publishing a payload with device-scope relaxed stores and the flag
// that announces it; device-scope release/acquire would instead write back /
// invalidate the whole L2 of the XCD, which is far too costly per task.
__device__ __forceinline__ void wave_drain() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// A list in LDS may be reached through a generic pointer (flat_ instructions)
// and through an LDS pointer (ds_ instructions).  The two take different paths
// to the LDS and are not ordered with each other, so a flat store can land
// after a later ds_read of the same address (and a ds_write after a later
// flat load).  Where the access kind of LDS data changes, wait for every
// outstanding memory operation of the wave.
__device__ __forceinline__ void lds_flat_fence() { wave_drain(); }

// relaxed device-scope accesses: coherent across the XCDs' L2 caches
template <class T>
__device__ __forceinline__ T ld_dev(const T* p) {
  return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T, class V>
__device__ __forceinline__ void st_dev(T* p, V v) {
  __hip_atomic_store(p, (T)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// relaxed work-group-scope accesses for a wave's private scratch
template <class T>
__device__ __forceinline__ T ld_wg(const T* p) {
  return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <class T, class V>
__device__ __forceinline__ void st_wg(T* p, V v) {
  __hip_atomic_store(p, (T)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <class T, class V>
__device__ __forceinline__ T add_wg(T* p, V v) {
  return __hip_atomic_fetch_add(p, (T)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <class T, class V>
__device__ __forceinline__ T or_wg(T* p, V v) {
  return __hip_atomic_fetch_or(p, (T)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ uint64_t rotl64(uint64_t x, int r) {
  return (x << r) | (x >> (64 - r));
}

}  // namespace mbe
