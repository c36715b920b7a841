// Learned scheduler (placeholder until the MLP kernels land).
#include "engine.hpp"

extern "C" fc_status fc_sched_trace(fc_graph*, const int32_t*, const int32_t*, int32_t, fc_trace_row*, int32_t,
                                    int32_t*) {
  return fc::fail(FC_ERR_UNSUPPORTED, "not built yet");
}
extern "C" fc_status fc_sched_train(fc_graph*, const fc_trace_row*, int32_t, int32_t, float, uint64_t, float*) {
  return fc::fail(FC_ERR_UNSUPPORTED, "not built yet");
}
extern "C" fc_status fc_sched_set_params(fc_graph*, const float*, int32_t) { return fc::fail(FC_ERR_UNSUPPORTED, "not built yet"); }
extern "C" fc_status fc_sched_get_params(fc_graph*, float*, int32_t) { return fc::fail(FC_ERR_UNSUPPORTED, "not built yet"); }
extern "C" fc_status fc_sched_forward(fc_graph*, const float*, int32_t, float*) { return fc::fail(FC_ERR_UNSUPPORTED, "not built yet"); }
extern "C" fc_status fc_sched_gradient(fc_graph*, const float*, const float*, int32_t, float*, float*) {
  return fc::fail(FC_ERR_UNSUPPORTED, "not built yet");
}
