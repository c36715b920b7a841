// Batched bit-parallel s–t waves (placeholder until the wave kernels land).
#include "engine.hpp"

namespace fcd {
struct BatchWorkspace {};
void batch_free(BatchWorkspace* b) { delete b; }
}  // namespace fcd

extern "C" fc_status fc_stcon_batch(fc_graph* g, const int32_t* s, const int32_t* t, int32_t n, int32_t* connected,
                                    int32_t* dist, const fc_batch_opts* opts, fc_batch_stats* stats) {
  return fc::fail(FC_ERR_UNSUPPORTED, "fc_stcon_batch: not built yet");
}
