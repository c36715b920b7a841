// Internal definition of fc_graph: the device copy of one CSR graph, its HIP
// stream, and all per-query workspace, allocated once at creation.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "common.hpp"
#include "fastcon.h"
#include "visit.hpp"

namespace fcd {
struct BatchWorkspace;
}

struct fc_graph {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int64_t V = 0, E = 0;
  bool off64 = false;
  void* d_off = nullptr;
  int32_t* d_col = nullptr;
  int cus = 0;
  int blocks = 0;               // cooperative grid of the visit kernel
  int32_t hub_threshold = 4096;
  uint32_t hcap = 1;
  uint32_t qcap = 1;

  // Visit workspace: one allocation, so a step-by-step trace can snapshot it.
  char* d_ws = nullptr;
  size_t ws_bytes = 0;
  char* d_snap = nullptr;       // lazily allocated copy used by the tracer
  int32_t* d_state = nullptr;
  int32_t* d_parent = nullptr;
  uint32_t* d_visited = nullptr;
  uint32_t* d_fbm[2][3] = {};
  int32_t* d_q[2][2] = {};
  int32_t* d_hq[2][2] = {};
  int64_t* d_hb[2][2] = {};
  fcd::VisitCtl* d_ctl = nullptr;

  fcd::VisitOut* h_out = nullptr;   // pinned, out_cap entries
  int32_t out_cap = 0;

  bool parents_valid = false;
  bool levels_valid = false;
  float* d_mlp = nullptr;           // scheduler parameters (device)
  bool mlp_set = false;
  std::vector<float> h_mlp;

  fcd::BatchWorkspace* batch = nullptr;
  uint64_t spin_limit = 0;          // 100 MHz ticks
};

namespace fcd {

fc_status hip_fail(hipError_t e, const char* what);
fc_status ensure_out(fc_graph* g, int32_t n);

// Launch one visit (cooperative) writing into g->h_out[slot]; no sync.
fc_status launch_one(fc_graph* g, int mode, int policy, int32_t src, int32_t dst, bool parents, int slot,
                     int force_side = -1, int force_strategy = -1, int max_steps = 0, bool resume = false,
                     int lvl0 = 0, int lvl1 = 0, float* trace_feat = nullptr);

template <typename OffT>
hipError_t launch_visit(const VisitParams<OffT>& p, int blocks, hipStream_t s);
template <typename OffT>
int visit_occupancy();

void batch_free(BatchWorkspace* b);

}  // namespace fcd

#define FC_HIP(call)                                              \
  do {                                                            \
    hipError_t _e = (call);                                       \
    if (_e != hipSuccess) return fcd::hip_fail(_e, #call);        \
  } while (0)
