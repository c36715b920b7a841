// Wave64 primitives for gfx950.
//
// These replace the reference's 32-lane PTX scans and lane masks
// (BFS-4K/src/device/Util/prefixSumAsm.cu:5-139, BFS-4K/src/device/Util/ptx.cu:6-85),
// and its warp-aggregated frontier reservation (BFS-4K/src/device/Util/GlobalWrite.cu:13-46).
// On CDNA a wave has 64 lanes, so ballots are 64-bit and ranks come from
// v_mbcnt_lo/hi. The reference's software grid barrier
// (BFS-4K/src/device/Util/GlobalSync.cu:5-36) has no counterpart: levels are
// separated by kernel launches (see visit.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fcd {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Number of set bits of m strictly below this lane.
__device__ __forceinline__ uint32_t lane_rank(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return (uint64_t)__ballot(p ? 1 : 0); }

// Reserve one slot per lane with `want` using a single atomic per wave. Must
// be called by all active lanes of the wave with the same counter; the first
// active lane issues the atomic.
__device__ __forceinline__ uint32_t wave_reserve(uint32_t* counter, bool want, uint64_t* mask_out) {
  const uint64_t act = ballot(true);
  const uint64_t m = ballot(want);
  *mask_out = m;
  uint32_t base = 0;
  if (m) {
    const int leader = __ffsll((unsigned long long)act) - 1;
    if (lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = (uint32_t)__shfl((int)base, leader);
  }
  return base + lane_rank(m);
}

// v of lane src, for a src that every lane holds alike (a ballot's bit, a
// constant): one scalar read of the lane's register instead of an LDS
// permute. The lane need not be active.
__device__ __forceinline__ int32_t bcast(int32_t v, int src) { return __builtin_amdgcn_readlane(v, src); }
__device__ __forceinline__ int64_t bcast64(int64_t v, int src) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(uint64_t)v, src);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)v >> 32), src);
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ int64_t shfl64(int64_t v, int src) {
  int lo = __shfl((int)(uint32_t)(uint64_t)v, src);
  int hi = __shfl((int)(uint32_t)((uint64_t)v >> 32), src);
  return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// Sum, OR or minimum of a 64-bit value over a wave whose 64 lanes are all
// active, with DPP row shifts and row broadcasts (VALU only, no LDS): after
// the shifts by 1, 2, 4 and 8 lane 15 of each row of 16 holds the row's
// total, the broadcast of lane 15 folds rows 0 and 2 into rows 1 and 3, and
// the broadcast of lane 31 folds rows 0-1 into row 3. Lane 63 then holds the
// wave's total, which every lane reads. Lanes that a step has no source for
// take the operation's identity.
enum WaveOp : int { kWaveSum, kWaveOr, kWaveMin };

template <int Op>
__device__ __forceinline__ uint64_t wave_apply(uint64_t a, uint64_t b) {
  return Op == kWaveSum ? a + b : Op == kWaveOr ? (a | b) : (b < a ? b : a);
}

template <int Op, int Ctrl, int RowMask>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t v) {
  // the identity: 0 written for lanes without a source (bound_ctrl), or the
  // all-ones old value kept for them
  constexpr bool kOnes = Op == kWaveMin;
  const int old = kOnes ? -1 : 0;
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(old, (int)(uint32_t)v, Ctrl, RowMask, 0xf, !kOnes);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(old, (int)(uint32_t)(v >> 32), Ctrl, RowMask, 0xf, !kOnes);
  return ((uint64_t)hi << 32) | lo;
}

template <int Op>
__device__ __forceinline__ uint64_t wave_reduce_full_u64(uint64_t v) {
  constexpr int kRowShr = 0x110, kRowBcast15 = 0x142, kRowBcast31 = 0x143;
  v = wave_apply<Op>(v, dpp_u64<Op, kRowShr + 1, 0xf>(v));
  v = wave_apply<Op>(v, dpp_u64<Op, kRowShr + 2, 0xf>(v));
  v = wave_apply<Op>(v, dpp_u64<Op, kRowShr + 4, 0xf>(v));
  v = wave_apply<Op>(v, dpp_u64<Op, kRowShr + 8, 0xf>(v));
  v = wave_apply<Op>(v, dpp_u64<Op, kRowBcast15, 0xa>(v));
  v = wave_apply<Op>(v, dpp_u64<Op, kRowBcast31, 0xc>(v));
  return (uint64_t)bcast64((int64_t)v, 63);
}

// The same, without any work when every lane holds the identity.
template <int Op>
__device__ __forceinline__ uint64_t wave_reduce_full_nz_u64(uint64_t v) {
  constexpr uint64_t id = Op == kWaveMin ? ~0ull : 0ull;
  return ballot(v != id) ? wave_reduce_full_u64<Op>(v) : id;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
  for (int o = 32; o > 0; o >>= 1) {
    uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, o);
    uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), o);
    v += ((uint64_t)hi << 32) | lo;
  }
  return v;
}

template <typename T>
__device__ __forceinline__ T ld_agent(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void st_agent(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace fcd
