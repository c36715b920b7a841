// Internal definition of fc_graph: the device copy of one CSR graph, its HIP
// stream, and all per-query workspace, allocated once at creation.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
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
  float f_branch = 0.f;         // scheduler feature: log2(1 + sum d^2 / sum d) / 16
  uint32_t hcap = 1;
  uint32_t qcap = 1;

  // Visit workspace: one allocation.
  char* d_ws = nullptr;
  size_t ws_bytes = 0;
  int32_t* d_state = nullptr;
  unsigned long long* d_sst = nullptr;  // s–t state words: (epoch << 32) | colour | level
  uint32_t epoch_next = 1;              // epoch of the next s–t query (0 = never visited)
  int32_t* d_parent = nullptr;
  uint32_t* d_visited = nullptr;
  uint32_t* d_fbm[2][3] = {};
  int32_t* d_q[2][2] = {};
  int32_t* d_hq[2][2] = {};
  int64_t* d_hb[2][2] = {};
  fcd::VisitCtl* d_ctl = nullptr;

  // Query lists and results of the phase machine (device), with a pinned
  // host copy of the results; capacity q_cap.
  int32_t* d_qs = nullptr;
  int32_t* d_qt = nullptr;
  fcd::VisitOut* d_out = nullptr;
  fcd::VisitOut* h_out = nullptr;
  int32_t q_cap = 0;
  int32_t* h_phase = nullptr;       // pinned: phase + qi read back between batches
  // Captured hipGraph of kGraphLaunches identical phase launches.
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  double est_launches[2] = {10.0, 3.0};  // working launches per query (running estimate, per mode)
  int32_t vw_fixed = 0;
  bool use_graphs = true;           // replay captured launches (FASTCON_PLAIN_GRAPHS=0: direct launches)

  // Degree relabelling (absent with FC_GRAPH_KEEP_IDS): the device graph
  // uses new ids; perm maps a caller id to its device id, iperm back.
  bool relabeled = false;
  bool lists_sorted = false;  // device adjacency lists in ascending id order (renumbered, not FC_GRAPH_KEEP_ORDER)
  int64_t v_active = 0;  // vertices [0, v_active) have degree > 0: the degree-ordered
                         // numbering puts every isolated vertex after them; V otherwise
  int32_t* d_perm = nullptr;
  int32_t* d_iperm = nullptr;
  std::vector<int32_t> h_perm, h_iperm;
  // The last BFS's levels and parents in caller ids (relabelled graphs).
  int32_t* d_lv_out = nullptr;
  int32_t* d_par_out = nullptr;

  bool parents_valid = false;
  bool levels_valid = false;
  float* d_mlp = nullptr;           // scheduler parameters (device): canonical, then in mlp_t_index order
  bool mlp_set = false;
  std::vector<float> h_mlp;

  fcd::BatchWorkspace* batch = nullptr;
};

namespace fcd {

constexpr int kGraphLaunches = 16;
constexpr int64_t kPollLaunches = 4096;  // launches queued between two host polls of the phase word

fc_status hip_fail(hipError_t e, const char* what);
fc_status ensure_queries(fc_graph* g, int32_t n);

// Run n visits through the phase machine (s[i], t[i]; t = -1 for BFS) and
// leave their results in g->h_out[0..n). *ms receives the device time.
fc_status run_visits(fc_graph* g, int mode, const int32_t* s, const int32_t* t, int32_t n, bool parents,
                     double* ms);

// One launch of the visit kernel on the control block ctl (which carries the
// launch parameters); launches alternate parity 0, 1, 0, ...
template <typename OffT>
hipError_t launch_phase(VisitCtl* ctl, int parity, int blocks, hipStream_t s);

// Control block for a run of n queries: s–t queries start directly in their
// level phase (query i has epoch epoch0 + i), a BFS starts with its reset.
template <typename OffT>
inline VisitCtl visit_init(const VisitParams<OffT>& prm, int mode, int32_t n, const int32_t* s, const int32_t* t,
                           uint32_t epoch0, int vw_fixed, bool parents) {
  VisitCtl c;
  std::memset(&c, 0, sizeof(c));
  std::memcpy(c.params, &prm, sizeof(prm));
  VisitState& st = c.st[0];
  st.phase = mode == kModeBfs ? kPhaseReset : kPhaseLevel;
  st.qi = 0;
  st.src = s[0];
  st.dst = t[0];
  st.nsrc = st.ndst = -1;
  st.epoch = epoch0;
  c.best[0] = c.best[1] = ~0ull;
  c.nq = n;
  c.mode = mode;
  c.vw_fixed = vw_fixed;
  c.record_parents = parents ? 1 : 0;
  return c;
}
template <typename OffT>
int phase_occupancy();

void batch_free(BatchWorkspace* b);

}  // namespace fcd

#define FC_HIP(call)                                              \
  do {                                                            \
    hipError_t _e = (call);                                       \
    if (_e != hipSuccess) return fcd::hip_fail(_e, #call);        \
  } while (0)
