// Wave64 primitives for gfx950 and the bounded grid barrier used by the
// persistent level-loop kernels.
//
// These replace the reference's 32-lane PTX scans and lane masks
// (BFS-4K/src/device/Util/prefixSumAsm.cu:5-139, BFS-4K/src/device/Util/ptx.cu:6-85),
// its warp-aggregated frontier reservation (BFS-4K/src/device/Util/GlobalWrite.cu:13-46)
// and its software grid barrier (BFS-4K/src/device/Util/GlobalSync.cu:5-36,
// which spun on cub::ThreadLoad<LOAD_CG>). On CDNA a wave has 64 lanes, so
// ballots are 64-bit and ranks come from v_mbcnt_lo/hi.
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

__device__ __forceinline__ int64_t shfl64(int64_t v, int src) {
  int lo = __shfl((int)(uint32_t)(uint64_t)v, src);
  int hi = __shfl((int)(uint32_t)((uint64_t)v >> 32), src);
  return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
  for (int o = 32; o > 0; o >>= 1) {
    uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, o);
    uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), o);
    v += ((uint64_t)hi << 32) | lo;
  }
  return v;
}

__device__ __forceinline__ uint64_t wave_or_u64(uint64_t v) {
  for (int o = 32; o > 0; o >>= 1) {
    uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, o);
    uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), o);
    v |= ((uint64_t)hi << 32) | lo;
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

// Grid barrier state in device memory; zeroed before each launch.
struct GridBar {
  uint32_t count;
  uint32_t gen;
  uint32_t abort;
  uint32_t pad;
};

// Sense-counting grid barrier for a cooperative launch (all blocks resident).
// Each wait is bounded by `limit` ticks of the 100 MHz real-time counter; on
// expiry, or when another block already gave up, it sets `abort` and returns
// false so every caller leaves its level loop and the grid drains.
// Release/acquire fences at agent scope make the plain stores of one phase
// visible to the plain loads of the next across XCDs. Loop-exit decisions
// taken after the barrier read only values completed before it, so every
// block takes the same decision.
__device__ __forceinline__ bool grid_sync(GridBar* b, uint32_t nblocks, uint64_t limit) {
  __shared__ int s_ok;
  // Every wave first waits for its own outstanding memory operations,
  // including non-returning atomics (atomicMin/atomicAdd whose result is
  // unused), so that the block's arrival below publishes all of them.
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    int ok = 1;
    const uint32_t g = __hip_atomic_load(&b->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const uint32_t arrived =
        __hip_atomic_fetch_add(&b->count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (arrived == nblocks - 1) {
      __hip_atomic_store(&b->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&b->gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      const uint64_t t0 = wall_clock64();
      while (__hip_atomic_load(&b->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
        if (__hip_atomic_load(&b->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
          ok = 0;
          break;
        }
        if (wall_clock64() - t0 > limit) {
          __hip_atomic_store(&b->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          ok = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (__hip_atomic_load(&b->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) ok = 0;
    s_ok = ok;
  }
  __syncthreads();
  return s_ok != 0;
}

}  // namespace fcd
