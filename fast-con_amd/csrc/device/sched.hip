// Learned scheduler (placeholder until the MLP kernels land).
#include "engine.hpp"

#include <algorithm>
#include <cmath>

extern "C" fc_status fc_sched_trace(fc_graph*, const int32_t*, const int32_t*, int32_t, fc_trace_row*, int32_t,
                                    int32_t*) {
  return fc::fail(FC_ERR_UNSUPPORTED, "not built yet");
}
extern "C" fc_status fc_sched_train(fc_graph*, const fc_trace_row*, int32_t, int32_t, float, uint64_t, float*) {
  return fc::fail(FC_ERR_UNSUPPORTED, "not built yet");
}
// Parameters live on the device (read by the batch kernel's learned policy)
// with a host copy. Passing params == NULL and n == 0 removes them, which
// makes the learned policy unavailable again.
extern "C" fc_status fc_sched_set_params(fc_graph* g, const float* params, int32_t n) {
  fc::clear_error();
  if (!g) return fc::fail(FC_ERR_INVALID, "fc_sched_set_params: null graph");
  if (!params && n == 0) {
    g->mlp_set = false;
    return FC_OK;
  }
  if (!params || n != fcd::kMlpParams)
    return fc::fail(FC_ERR_INVALID, "fc_sched_set_params: expected %d parameters", fcd::kMlpParams);
  for (int32_t i = 0; i < n; ++i)
    if (!std::isfinite(params[i])) return fc::fail(FC_ERR_INVALID, "fc_sched_set_params: parameter %d is not finite", i);
  g->h_mlp.assign(params, params + n);
  FC_HIP(hipMemcpyAsync(g->d_mlp, g->h_mlp.data(), sizeof(float) * (size_t)n, hipMemcpyHostToDevice, g->stream));
  FC_HIP(hipStreamSynchronize(g->stream));
  g->mlp_set = true;
  return FC_OK;
}

extern "C" fc_status fc_sched_get_params(fc_graph* g, float* params, int32_t n) {
  fc::clear_error();
  if (!g || !params) return fc::fail(FC_ERR_INVALID, "fc_sched_get_params: null argument");
  if (n != fcd::kMlpParams)
    return fc::fail(FC_ERR_INVALID, "fc_sched_get_params: expected %d parameters", fcd::kMlpParams);
  if (!g->mlp_set) return fc::fail(FC_ERR_INVALID, "fc_sched_get_params: no scheduler parameters are set");
  std::copy(g->h_mlp.begin(), g->h_mlp.end(), params);
  return FC_OK;
}
extern "C" fc_status fc_sched_forward(fc_graph*, const float*, int32_t, float*) { return fc::fail(FC_ERR_UNSUPPORTED, "not built yet"); }
extern "C" fc_status fc_sched_gradient(fc_graph*, const float*, const float*, int32_t, float*, float*) {
  return fc::fail(FC_ERR_UNSUPPORTED, "not built yet");
}
