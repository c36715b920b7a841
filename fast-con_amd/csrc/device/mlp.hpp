// Scheduler MLP evaluation in the batch kernel's finalising wave.
#pragma once

#include "visit.hpp"

namespace fcd {

// Scheduler MLP (12 -> 32 -> 32 -> 9, ReLU); the output is a predicted
// log-cost per action. Canonical parameter layout (fc_sched_set_params):
// W1[32][12] b1[32] W2[32][32] b2[32] W3[9][32] b3[9].

// Position of canonical parameter i in the wave-friendly layout used by
// mlp_pick_wave_t: W1, W2 and W3 transposed (input-major), so that the lanes
// of one step read consecutive words; biases unchanged.
__host__ __device__ inline int mlp_t_index(int i) {
  constexpr int oB1 = kHidden * kFeatures, oW2 = oB1 + kHidden, oB2 = oW2 + kHidden * kHidden;
  constexpr int oW3 = oB2 + kHidden, oB3 = oW3 + kActions * kHidden;
  if (i < oB1) return (i % kFeatures) * kHidden + i / kFeatures;
  if (i < oW2 || (i >= oB2 && i < oW3) || i >= oB3) return i;
  if (i < oB2) {
    const int r = i - oW2;
    return oW2 + (r % kHidden) * kHidden + r / kHidden;
  }
  const int r = i - oW3;
  return oW3 + (r % kHidden) * kActions + r / kHidden;
}

// Learned choice with a fallback, evaluated by one wave on weights in the
// mlp_t_index layout (in LDS); all lanes call with identical x and heur.
// Lane j computes hidden unit j, summing inputs in ascending order (as the
// training kernels do), and the activations are broadcast by lane reads.
// The network's best action replaces the heuristic's action `heur` only
// when its predicted log-cost is lower by at least `margin`.
__device__ inline int mlp_pick_wave_t(const float* wt, const float* x, int heur, float margin) {
  constexpr int oB1 = kHidden * kFeatures, oW2 = oB1 + kHidden, oB2 = oW2 + kHidden * kHidden;
  constexpr int oW3 = oB2 + kHidden, oB3 = oW3 + kActions * kHidden;
  const int j = (int)__lane_id() & (kHidden - 1);
  float s = wt[oB1 + j];
#pragma unroll
  for (int i = 0; i < kFeatures; ++i) s = __fmaf_rn(wt[i * kHidden + j], x[i], s);
  const float h1 = s > 0.f ? s : 0.f;
  s = wt[oB2 + j];
#pragma unroll
  for (int i = 0; i < kHidden; ++i)
    s = __fmaf_rn(wt[oW2 + i * kHidden + j], __int_as_float(__builtin_amdgcn_readlane(__float_as_int(h1), i)), s);
  const float h2 = s > 0.f ? s : 0.f;
  const int a = (int)__lane_id() < kActions ? (int)__lane_id() : 0;
  s = wt[oB3 + a];
#pragma unroll
  for (int i = 0; i < kHidden; ++i)
    s = __fmaf_rn(wt[oW3 + i * kActions + a], __int_as_float(__builtin_amdgcn_readlane(__float_as_int(h2), i)), s);
  int best = 0;
  float bv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(s), 0));
#pragma unroll
  for (int k = 1; k < kActions; ++k) {
    const float v = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(s), k));
    if (v < bv) {
      bv = v;
      best = k;
    }
  }
  const float hv = __shfl(s, heur);
  return bv + margin < hv ? best : heur;
}

}  // namespace fcd
