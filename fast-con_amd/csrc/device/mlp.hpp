// Scheduler MLP evaluation shared by the visit and batch kernels.
#pragma once

#include "visit.hpp"

namespace fcd {

// Scheduler MLP (12 -> 32 -> 32 -> 6, ReLU), evaluated by one thread; the
// output is a predicted log-cost per action. Parameter layout:
// W1[32][12] b1[32] W2[32][32] b2[32] W3[6][32] b3[6].
__device__ inline int mlp_argmin(const float* w, const float* x) {
  const float* W1 = w;
  const float* b1 = W1 + kHidden * kFeatures;
  const float* W2 = b1 + kHidden;
  const float* b2 = W2 + kHidden * kHidden;
  const float* W3 = b2 + kHidden;
  const float* b3 = W3 + kActions * kHidden;
  float h1[kHidden], h2[kHidden];
  for (int j = 0; j < kHidden; ++j) {
    float s = b1[j];
    for (int i = 0; i < kFeatures; ++i) s = __fmaf_rn(W1[j * kFeatures + i], x[i], s);
    h1[j] = s > 0.f ? s : 0.f;
  }
  for (int j = 0; j < kHidden; ++j) {
    float s = b2[j];
    for (int i = 0; i < kHidden; ++i) s = __fmaf_rn(W2[j * kHidden + i], h1[i], s);
    h2[j] = s > 0.f ? s : 0.f;
  }
  int best = 0;
  float bv = 0.f;
  for (int a = 0; a < kActions; ++a) {
    float s = b3[a];
    for (int i = 0; i < kHidden; ++i) s = __fmaf_rn(W3[a * kHidden + i], h2[i], s);
    if (a == 0 || s < bv) {
      bv = s;
      best = a;
    }
  }
  return best;
}

// The same network evaluated by one wave (all 64 lanes call, x identical in
// every lane): lane j computes hidden unit j with the same summation order
// as mlp_argmin, so both give the same action.
__device__ inline int mlp_argmin_wave(const float* w, const float* x) {
  const float* W1 = w;
  const float* b1 = W1 + kHidden * kFeatures;
  const float* W2 = b1 + kHidden;
  const float* b2 = W2 + kHidden * kHidden;
  const float* W3 = b2 + kHidden;
  const float* b3 = W3 + kActions * kHidden;
  const int j = (int)__lane_id() & (kHidden - 1);
  float s = b1[j];
  for (int i = 0; i < kFeatures; ++i) s = __fmaf_rn(W1[j * kFeatures + i], x[i], s);
  const float h1 = s > 0.f ? s : 0.f;
  s = b2[j];
  for (int i = 0; i < kHidden; ++i) s = __fmaf_rn(W2[j * kHidden + i], __shfl(h1, i), s);
  const float h2 = s > 0.f ? s : 0.f;
  const int a = (int)__lane_id() < kActions ? (int)__lane_id() : 0;
  s = b3[a];
  for (int i = 0; i < kHidden; ++i) s = __fmaf_rn(W3[a * kHidden + i], __shfl(h2, i), s);
  int best = 0;
  float bv = __shfl(s, 0);
  for (int k = 1; k < kActions; ++k) {
    const float v = __shfl(s, k);
    if (v < bv) {
      bv = v;
      best = k;
    }
  }
  return best;
}

// Learned choice with a fallback: the network's best action replaces the
// heuristic's action `heur` only when its predicted log-cost is lower by at
// least `margin`. All lanes call with identical x and heur.
__device__ inline int mlp_pick_wave(const float* w, const float* x, int heur, float margin) {
  const float* W1 = w;
  const float* b1 = W1 + kHidden * kFeatures;
  const float* W2 = b1 + kHidden;
  const float* b2 = W2 + kHidden * kHidden;
  const float* W3 = b2 + kHidden;
  const float* b3 = W3 + kActions * kHidden;
  const int j = (int)__lane_id() & (kHidden - 1);
  float s = b1[j];
  for (int i = 0; i < kFeatures; ++i) s = __fmaf_rn(W1[j * kFeatures + i], x[i], s);
  const float h1 = s > 0.f ? s : 0.f;
  s = b2[j];
  for (int i = 0; i < kHidden; ++i) s = __fmaf_rn(W2[j * kHidden + i], __shfl(h1, i), s);
  const float h2 = s > 0.f ? s : 0.f;
  const int a = (int)__lane_id() < kActions ? (int)__lane_id() : 0;
  s = b3[a];
  for (int i = 0; i < kHidden; ++i) s = __fmaf_rn(W3[a * kHidden + i], __shfl(h2, i), s);
  int best = 0;
  float bv = __shfl(s, 0);
  for (int k = 1; k < kActions; ++k) {
    const float v = __shfl(s, k);
    if (v < bv) {
      bv = v;
      best = k;
    }
  }
  const float hv = __shfl(s, heur);
  return bv + margin < hv ? best : heur;
}

}  // namespace fcd
