// LDS staging of hub-queue appends, shared by the visit and batch kernels.
//
// A hub entry is a vertex of degree >= hub_threshold together with the base
// of its edges in the concatenated adjacency of the hub queue. Appending it
// with its own 64-bit atomicAdd on the single (entries << 40 | edges) counter
// serialises on that word when a level reaches thousands of hubs, so entries
// are staged per wave in LDS and the block publishes them with one atomicAdd,
// deriving every entry's edge base from prefix sums of the degrees.
#pragma once

#include "visit.hpp"

namespace fcd {

constexpr uint32_t kHubStage = 64;  // staged hub entries per wave (one per lane)

struct HubStage {
  int32_t* u;     // LDS, kHubStage entries
  int32_t* d;     // LDS, their degrees
  uint32_t cnt;   // wave-uniform
};

struct HubShared {
  uint32_t cnt[kWavesPerBlock];
  unsigned long long edges[kWavesPerBlock];
  unsigned long long base;
};

// Inclusive wave scan of 64-bit values.
__device__ __forceinline__ unsigned long long wave_incl_scan_u64(unsigned long long v) {
  const int lane = lane_id();
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t lo = (uint32_t)__shfl_up((int)(uint32_t)v, o);
    const uint32_t hi = (uint32_t)__shfl_up((int)(uint32_t)(v >> 32), o);
    if (lane >= o) v += ((unsigned long long)hi << 32) | lo;
  }
  return v;
}

// Write this wave's staged entries at (slot0, edge0) onwards.
__device__ __forceinline__ void hub_write(HubStage& h, uint32_t slot0, unsigned long long edge0, int32_t* hq,
                                          int64_t* hb, uint32_t hcap, uint32_t* err) {
  const int lane = lane_id();
  const unsigned long long d = (uint32_t)lane < h.cnt ? (unsigned long long)h.d[lane] : 0ull;
  const unsigned long long incl = wave_incl_scan_u64(d);
  if ((uint32_t)lane < h.cnt) {
    const uint32_t slot = slot0 + (uint32_t)lane;
    if (slot < hcap) {
      hq[slot] = h.u[lane];
      hb[slot] = (int64_t)(edge0 + incl - d);
    } else {
      atomicOr(err, 1u);
    }
  }
}

// Wave-local flush (stage full): one atomic for this wave.
__device__ __forceinline__ void hub_wave_flush(HubStage& h, unsigned long long* counter, int32_t* hq, int64_t* hb,
                                               uint32_t hcap, uint32_t* err) {
  if (!h.cnt) return;
  __builtin_amdgcn_wave_barrier();
  const int lane = lane_id();
  const unsigned long long d = (uint32_t)lane < h.cnt ? (unsigned long long)h.d[lane] : 0ull;
  const unsigned long long tot = wave_sum_u64(d);
  unsigned long long old = 0;
  if (lane == 0) old = atomicAdd(counter, ((unsigned long long)h.cnt << kHubEdgeBits) | tot);
  old = (unsigned long long)shfl64((int64_t)old, 0);
  hub_write(h, (uint32_t)(old >> kHubEdgeBits), old & kHubEdgeMask, hq, hb, hcap, err);
  __builtin_amdgcn_wave_barrier();
  h.cnt = 0;
}

// Wave-collective append of (u, deg) for lanes with want.
__device__ __forceinline__ void hub_push(HubStage& h, bool want, int32_t u, int64_t deg, unsigned long long* counter,
                                         int32_t* hq, int64_t* hb, uint32_t hcap, uint32_t* err) {
  const uint64_t m = ballot(want);
  if (!m) return;
  const uint32_t k = (uint32_t)__popcll(m);
  if (h.cnt + k > kHubStage) hub_wave_flush(h, counter, hq, hb, hcap, err);
  if (want) {
    const uint32_t i = h.cnt + lane_rank(m);
    h.u[i] = u;
    h.d[i] = (int32_t)deg;
  }
  h.cnt += k;
}

// Block-collective publication (all threads call, uniform control flow).
__device__ __forceinline__ void hub_block_flush(HubStage& h, unsigned long long* counter, int32_t* hq, int64_t* hb,
                                                uint32_t hcap, uint32_t* err, HubShared& sh) {
  const int w = (int)(threadIdx.x >> 6);
  const int lane = lane_id();
  const unsigned long long d = (uint32_t)lane < h.cnt ? (unsigned long long)h.d[lane] : 0ull;
  const unsigned long long tot = wave_sum_u64(d);
  if (lane == 0) {
    sh.cnt[w] = h.cnt;
    sh.edges[w] = tot;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t c = 0;
    unsigned long long e = 0;
    for (int i = 0; i < kWavesPerBlock; ++i) {
      const uint32_t ci = sh.cnt[i];
      const unsigned long long ei = sh.edges[i];
      sh.cnt[i] = c;
      sh.edges[i] = e;
      c += ci;
      e += ei;
    }
    sh.base = c ? atomicAdd(counter, ((unsigned long long)c << kHubEdgeBits) | e) : 0ull;
  }
  __syncthreads();
  if (h.cnt) {
    const unsigned long long base = sh.base;
    hub_write(h, (uint32_t)(base >> kHubEdgeBits) + sh.cnt[w], (base & kHubEdgeMask) + sh.edges[w], hq, hb, hcap,
              err);
  }
  h.cnt = 0;
  __syncthreads();
}

}  // namespace fcd
