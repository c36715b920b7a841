// C ABI of the device engine: graph upload, BFS and plain s–t queries.
//
// fc_graph_create_* replaces cudaGraph::cudaGraph (BFS-4K/src/device/cudaGraph.cu:7-36):
// the CSR is uploaded once and every per-query buffer is allocated up front,
// sized for the graph (queues of V entries per colour and parity, three
// frontier bitmaps per colour, a hub queue sized by the number of vertices of
// degree >= hub_threshold) instead of the reference's free/2 frontier buffers.
// fc_bfs replaces Reset + cudaBFS4K (BFS-4K/src/device/BFS_WorkEfficient.cu:81-133);
// fc_stcon replaces make_stcon's single-GPU path (st-conn/make_stcon.c:730-1497).
//
// Queries run through the device-side phase machine of visit.hip: the host
// writes the query list and an initial control block, then replays a
// captured hipGraph of identical phase launches (or launches them directly
// when only a few are needed), reading back the 8-byte phase word between
// batches to know whether every query has finished.
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

using fc::fail;

namespace fc {
fc_status validate_csr32(const int32_t* off, const int32_t* col, int64_t V, int64_t E, int64_t hub, int64_t* hubs);
fc_status validate_csr64(const int64_t* off, const int32_t* col, int64_t V, int64_t E, int64_t hub, int64_t* hubs);
void hub_first32(const int32_t* off, const int32_t* col, int64_t V, int32_t* out);
void hub_first64(const int64_t* off, const int32_t* col, int64_t V, int32_t* out);
void relabel32(const int32_t* off, const int32_t* col, int64_t V, bool keep_order, int32_t* perm, int32_t* iperm,
               int32_t* noff, int32_t* ncol);
void relabel64(const int64_t* off, const int32_t* col, int64_t V, bool keep_order, int32_t* perm, int32_t* iperm,
               int64_t* noff, int32_t* ncol);
}  // namespace fc

namespace fcd {

fc_status hip_fail(hipError_t e, const char* what) {
  return fail(e == hipErrorNoDevice || e == hipErrorNoBinaryForGpu || e == hipErrorInvalidDevice ? FC_ERR_NOGPU
                                                                                                  : FC_ERR_HIP,
              "%s: %s", what, hipGetErrorString(e));
}

fc_status ensure_queries(fc_graph* g, int32_t n) {
  if (n <= g->q_cap) return FC_OK;
  // the captured launches take only the control block, whose parameters
  // are rewritten by every run, so the graph stays valid
  (void)hipFree(g->d_qs);
  (void)hipFree(g->d_qt);
  (void)hipFree(g->d_out);
  if (g->h_out) (void)hipHostFree(g->h_out);
  g->d_qs = g->d_qt = nullptr;
  g->d_out = nullptr;
  g->h_out = nullptr;
  g->q_cap = 0;
  const int32_t cap = std::max<int32_t>(n, 1024);
  FC_HIP(hipMalloc((void**)&g->d_qs, sizeof(int32_t) * (size_t)cap));
  FC_HIP(hipMalloc((void**)&g->d_qt, sizeof(int32_t) * (size_t)cap));
  FC_HIP(hipMalloc((void**)&g->d_out, sizeof(VisitOut) * (size_t)cap));
  FC_HIP(hipHostMalloc((void**)&g->h_out, sizeof(VisitOut) * (size_t)cap, hipHostMallocDefault));
  g->q_cap = cap;
  return FC_OK;
}

template <typename OffT>
static VisitParams<OffT> make_params(fc_graph* g) {
  VisitParams<OffT> p;
  std::memset(&p, 0, sizeof(p));
  set_param(p.off, (const OffT*)g->d_off);
  set_param(p.col, g->d_col);
  p.V = (int32_t)g->V;
  p.E = g->E;
  p.hub_threshold = g->hub_threshold;
  set_param(p.state, g->d_state);
  set_param(p.sst, g->d_sst);
  set_param(p.parent, g->d_parent);
  set_param(p.visited, g->d_visited);
  for (int c = 0; c < 2; ++c) {
    for (int k = 0; k < 3; ++k) set_param(p.fbm[c][k], g->d_fbm[c][k]);
    for (int k = 0; k < 2; ++k) {
      set_param(p.q[c][k], g->d_q[c][k]);
      set_param(p.hq[c][k], g->d_hq[c][k]);
      set_param(p.hb[c][k], g->d_hb[c][k]);
    }
  }
  p.qcap = g->qcap;
  p.hcap = g->hcap;
  set_param(p.ctl, g->d_ctl);
  set_param(p.qs, g->d_qs);
  set_param(p.qt, g->d_qt);
  set_param(p.out, g->d_out);
  p.lazy_state = g->relabeled ? 1 : 0;
  p.v_active = (int32_t)g->v_active;
  return p;
}

static hipError_t launch_raw(fc_graph* g, int parity) {
  if (g->off64) return launch_phase<int64_t>(g->d_ctl, parity, g->blocks, g->stream);
  return launch_phase<int32_t>(g->d_ctl, parity, g->blocks, g->stream);
}

static fc_status ensure_graph(fc_graph* g) {
  if (g->gexec) return FC_OK;
  FC_HIP(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
  hipError_t le = hipSuccess;
  for (int i = 0; i < kGraphLaunches && le == hipSuccess; ++i) le = launch_raw(g, i & 1);  // parities 0, 1, ...
  hipGraph_t gr = nullptr;
  hipError_t ce = hipStreamEndCapture(g->stream, &gr);
  if (le != hipSuccess) return hip_fail(le, "phase kernel capture");
  if (ce != hipSuccess) return hip_fail(ce, "hipStreamEndCapture");
  g->graph = gr;
  FC_HIP(hipGraphInstantiate(&g->gexec, g->graph, nullptr, nullptr, 0));
  return FC_OK;
}

fc_status run_visits(fc_graph* g, int mode, const int32_t* s, const int32_t* t, int32_t n, bool parents,
                     double* ms) {
  *ms = 0.0;
  if (n <= 0) return FC_OK;
  fc_status st = ensure_queries(g, n);
  if (st != FC_OK) return st;
  if (g->use_graphs) {
    st = ensure_graph(g);
    if (st != FC_OK) return st;
  }
  // s–t queries tag the state words with one epoch each (0 = never
  // visited); before the counter would wrap, the words are cleared.
  uint32_t epoch0 = 0;
  if (mode == kModePlainSt) {
    if ((uint64_t)g->epoch_next + (uint64_t)n >= 0xFFFFFFF0ull) {
      FC_HIP(hipMemsetAsync(g->d_sst, 0, sizeof(unsigned long long) * (size_t)g->V, g->stream));
      g->epoch_next = 1;
    }
    epoch0 = g->epoch_next;
    g->epoch_next += (uint32_t)n;
  }
  FC_HIP(hipMemcpyAsync(g->d_qs, s, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, g->stream));
  FC_HIP(hipMemcpyAsync(g->d_qt, t, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, g->stream));
  const VisitCtl init = g->off64 ? visit_init(make_params<int64_t>(g), mode, n, s, t, epoch0, g->vw_fixed, parents)
                                 : visit_init(make_params<int32_t>(g), mode, n, s, t, epoch0, g->vw_fixed, parents);
  FC_HIP(hipMemcpyAsync(g->d_ctl, &init, sizeof(init), hipMemcpyHostToDevice, g->stream));
  // The async copies above read pageable host memory: finish them before the
  // caller's arrays can go away, and before timing starts.
  FC_HIP(hipStreamSynchronize(g->stream));

  // Launch budget: the running estimate of working launches per query, plus
  // the launch that notices the last query's end. A single BFS needs no
  // extra margin: a short budget costs one more host poll, and each surplus
  // launch of the full grid a few us. At most kPollLaunches are queued
  // between two polls: a high-diameter batch needs hundreds of thousands of
  // launches, and queueing them all at once holds that many dispatches (and
  // a profiler's record of each) in flight in the runtime. A poll costs one
  // host round trip per kPollLaunches launches of at least 1.6 us each.
  // Launches alternate parity 0, 1, 0, ... (the state a launch reads), and
  // the captured graphs hold an even number of launches starting at 0.
  const double est = g->est_launches[mode];
  int64_t expected = (int64_t)std::ceil((double)n * est) + 1;
  int64_t launched = 0;
  FC_HIP(hipEventRecord(g->ev0, g->stream));
  for (int round = 0;; ++round) {
    int64_t want = expected - launched;
    if (want < 1) want = 1;
    if (want > kPollLaunches) want = kPollLaunches;
    if (want < kGraphLaunches || !g->use_graphs) {
      for (int64_t i = 0; i < want; ++i) {
        hipError_t e = launch_raw(g, (int)(launched & 1));
        if (e != hipSuccess) return hip_fail(e, "phase kernel launch");
        ++launched;
      }
    } else {
      if (launched & 1) {
        hipError_t e = launch_raw(g, 1);
        if (e != hipSuccess) return hip_fail(e, "phase kernel launch");
        ++launched;
        --want;
      }
      const int64_t ng = (want + kGraphLaunches - 1) / kGraphLaunches;
      for (int64_t i = 0; i < ng; ++i) FC_HIP(hipGraphLaunch(g->gexec, g->stream));
      launched += ng * kGraphLaunches;
    }
    FC_HIP(hipEventRecord(g->ev1, g->stream));
    // the state the next launch would read
    FC_HIP(hipMemcpyAsync(g->h_phase, &g->d_ctl->st[launched & 1], 2 * sizeof(int32_t), hipMemcpyDeviceToHost,
                          g->stream));
    FC_HIP(hipStreamSynchronize(g->stream));
    const int32_t phase = g->h_phase[0], qi = g->h_phase[1];
    if (phase == kPhaseIdle) break;
    // Every query stops within V+2 levels (device guard), so this bound is
    // never reached by a working visit.
    if (launched > (int64_t)n * (g->V + 4) + 64)
      return fail(FC_ERR_CAPACITY, "phase machine did not finish after %lld launches", (long long)launched);
    // Not finished. While the budget lasts, keep to it; past it, budget the
    // remaining queries with a wider margin.
    if (launched < expected) continue;
    const int64_t remaining = std::max<int64_t>(n - qi, 1);
    expected = launched + (int64_t)std::ceil((double)remaining * (est + 1.0) * (1.0 + round)) + 1;
  }
  float fms = 0.f;
  FC_HIP(hipEventElapsedTime(&fms, g->ev0, g->ev1));
  *ms = fms;
  FC_HIP(hipMemcpyAsync(g->h_out, g->d_out, sizeof(VisitOut) * (size_t)n, hipMemcpyDeviceToHost, g->stream));
  FC_HIP(hipStreamSynchronize(g->stream));
  double wl = 0;
  for (int32_t i = 0; i < n; ++i) wl += g->h_out[i].launches;
  g->est_launches[mode] = 0.5 * g->est_launches[mode] + 0.5 * (wl / n);
  return FC_OK;
}

}  // namespace fcd

using fcd::hip_fail;

template <typename OffT>
static fc_status graph_create(const OffT* off, const int32_t* col, int64_t V, int64_t E, uint32_t flags,
                              fc_graph** out) {
  fc::clear_error();
  if (!out || !off || (E > 0 && !col) || V < 1 || V > INT32_MAX - 64 || E < 0)
    return fail(FC_ERR_INVALID, "fc_graph_create: invalid arguments (V=%lld E=%lld)", (long long)V, (long long)E);
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(FC_ERR_NOGPU, "no HIP device available");
  fc_graph* g = new fc_graph();
  g->V = V;
  g->E = E;
  g->off64 = sizeof(OffT) == 8;
  // Degree from which a frontier vertex goes to the edge-balanced hub
  // queue. 256 measured best at R-MAT s22 (batch 2.09 ms, plain 32.0K q/s,
  // BFS 0.235 ms, against 2.11 ms, 29.2K q/s and 0.237 ms at 512 and
  // 2.40 ms, 30.0K q/s and 0.244 ms at 64).
  const char* ht = std::getenv("FASTCON_HUB_THRESHOLD");
  const bool ht_set = ht && std::atoi(ht) > 0;
  g->hub_threshold = ht_set ? std::atoi(ht) : 256;
  const char* vwf = std::getenv("FASTCON_VW");
  if (vwf && std::atoi(vwf) > 0) g->vw_fixed = std::atoi(vwf);
  // Direct launches instead of replayed hipGraphs, for profiler runs: under
  // rocprofv3's kernel trace a hipGraphLaunch of the visit's graph ended in
  // a host segmentation fault inside the runtime, while direct launches of
  // the same kernel trace normally.
  const char* pg = std::getenv("FASTCON_PLAIN_GRAPHS");
  if (pg && std::strcmp(pg, "0") == 0) g->use_graphs = false;
  int64_t hubs = 0;
  fc_status st = sizeof(OffT) == 8
                     ? fc::validate_csr64((const int64_t*)off, col, V, E, g->hub_threshold, &hubs)
                     : fc::validate_csr32((const int32_t*)off, col, V, E, g->hub_threshold, &hubs);
  if (st != FC_OK) {
    delete g;
    return st;
  }
  // Hub-queue counters pack (entries << 40) | edges into 64 bits: at most
  // 2^24 - 1 hub entries per frontier. The default threshold doubles until
  // the hubs fit; a threshold the caller set is kept, and the graph refused.
  const int64_t hub_cap = (int64_t)(1ll << (64 - fcd::kHubEdgeBits));
  while (!ht_set && hubs >= hub_cap && g->hub_threshold < (1 << 30)) {
    g->hub_threshold *= 2;
    hubs = 0;
    for (int64_t v = 0; v < V; ++v) hubs += (int64_t)off[v + 1] - (int64_t)off[v] >= g->hub_threshold;
  }
  if (hubs >= hub_cap) {
    const int32_t thr = g->hub_threshold;
    delete g;
    return fail(FC_ERR_CAPACITY, "fc_graph_create: %lld vertices of degree >= %d exceed the hub queue's %lld entries",
                (long long)hubs, thr, (long long)((1ll << (64 - fcd::kHubEdgeBits)) - 1));
  }
  auto bail = [&](fc_status s) {
    fc_graph_destroy(g);
    return s;
  };
  {
    double d2 = 0.0;
    for (int64_t v = 0; v < V; ++v) {
      const double d = (double)((int64_t)off[v + 1] - (int64_t)off[v]);
      d2 += d * d;
    }
    g->f_branch = (float)(std::log2(1.0 + (E > 0 ? d2 / (double)E : 0.0)) / 16.0);
  }
#define CK(call)                                            \
  do {                                                      \
    hipError_t _e = (call);                                 \
    if (_e != hipSuccess) return bail(hip_fail(_e, #call)); \
  } while (0)
  CK(hipGetDevice(&g->device));
  hipDeviceProp_t prop;
  CK(hipGetDeviceProperties(&prop, g->device));
  g->cus = prop.multiProcessorCount;
  CK(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
  CK(hipEventCreate(&g->ev0));
  CK(hipEventCreate(&g->ev1));
  CK(hipMalloc(&g->d_off, sizeof(OffT) * (size_t)(V + 1)));
  CK(hipMalloc((void**)&g->d_col, sizeof(int32_t) * (size_t)std::max<int64_t>(E, 1)));
  if (!(flags & FC_GRAPH_KEEP_IDS)) {
    g->relabeled = true;
    g->h_perm.resize((size_t)V);
    g->h_iperm.resize((size_t)V);
    std::vector<OffT> noff((size_t)V + 1);
    std::vector<int32_t> ncol((size_t)std::max<int64_t>(E, 1));
    if (sizeof(OffT) == 8)
      fc::relabel64((const int64_t*)off, col, V, (flags & FC_GRAPH_KEEP_ORDER) != 0, g->h_perm.data(),
                    g->h_iperm.data(), (int64_t*)noff.data(), ncol.data());
    else
      fc::relabel32((const int32_t*)off, col, V, (flags & FC_GRAPH_KEEP_ORDER) != 0, g->h_perm.data(),
                    g->h_iperm.data(), (int32_t*)noff.data(), ncol.data());
    CK(hipMalloc((void**)&g->d_perm, sizeof(int32_t) * (size_t)V));
    CK(hipMalloc((void**)&g->d_iperm, sizeof(int32_t) * (size_t)V));
    CK(hipMemcpyAsync(g->d_perm, g->h_perm.data(), sizeof(int32_t) * (size_t)V, hipMemcpyHostToDevice, g->stream));
    CK(hipMemcpyAsync(g->d_iperm, g->h_iperm.data(), sizeof(int32_t) * (size_t)V, hipMemcpyHostToDevice, g->stream));
    CK(hipMemcpyAsync(g->d_off, noff.data(), sizeof(OffT) * (size_t)(V + 1), hipMemcpyHostToDevice, g->stream));
    if (E > 0) CK(hipMemcpyAsync(g->d_col, ncol.data(), sizeof(int32_t) * (size_t)E, hipMemcpyHostToDevice, g->stream));
    // decreasing degree: the vertices of degree > 0 are a prefix
    int64_t va = V;
    while (va > 0 && noff[(size_t)va] == noff[(size_t)va - 1]) --va;
    for (int64_t v = 0; v < va; ++v)
      if (noff[(size_t)v + 1] == noff[(size_t)v])
        return bail(fail(FC_ERR_INVALID, "internal: degree order broken at vertex %lld", (long long)v));
    g->v_active = va;
    g->lists_sorted = !(flags & FC_GRAPH_KEEP_ORDER);
  } else {
    g->v_active = V;
    CK(hipMemcpyAsync(g->d_off, off, sizeof(OffT) * (size_t)(V + 1), hipMemcpyHostToDevice, g->stream));
  }
  if (E > 0 && !g->relabeled) {
    if (flags & FC_GRAPH_KEEP_ORDER) {
      CK(hipMemcpyAsync(g->d_col, col, sizeof(int32_t) * (size_t)E, hipMemcpyHostToDevice, g->stream));
    } else {
      std::vector<int32_t> ordered((size_t)E);
      if (sizeof(OffT) == 8)
        fc::hub_first64((const int64_t*)off, col, V, ordered.data());
      else
        fc::hub_first32((const int32_t*)off, col, V, ordered.data());
      CK(hipMemcpyAsync(g->d_col, ordered.data(), sizeof(int32_t) * (size_t)E, hipMemcpyHostToDevice, g->stream));
    }
  }

  // Workspace layout (256-byte aligned pieces of one allocation).
  const size_t nw64 = (size_t)((V + 63) / 64);
  const size_t bm_bytes = nw64 * 2 * sizeof(uint32_t);
  g->qcap = (uint32_t)V;
  g->hcap = (uint32_t)std::max<int64_t>(hubs, 1);
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  size_t off_state = 0;
  size_t off_sst = off_state + al(sizeof(int32_t) * (size_t)V);
  size_t off_parent = off_sst + al(sizeof(unsigned long long) * (size_t)V);
  size_t off_vis = off_parent + al(sizeof(int32_t) * (size_t)V);
  size_t off_fbm = off_vis + al(bm_bytes);
  size_t off_q = off_fbm + 6 * al(bm_bytes);
  size_t off_hq = off_q + 4 * al(sizeof(int32_t) * (size_t)g->qcap);
  size_t off_hb = off_hq + 4 * al(sizeof(int32_t) * (size_t)g->hcap);
  size_t off_ctl = off_hb + 4 * al(sizeof(int64_t) * (size_t)g->hcap);
  g->ws_bytes = off_ctl + al(sizeof(fcd::VisitCtl));
  CK(hipMalloc((void**)&g->d_ws, g->ws_bytes));
  CK(hipMemsetAsync(g->d_ws, 0, g->ws_bytes, g->stream));
  g->d_state = (int32_t*)(g->d_ws + off_state);
  g->d_sst = (unsigned long long*)(g->d_ws + off_sst);
  g->d_parent = (int32_t*)(g->d_ws + off_parent);
  g->d_visited = (uint32_t*)(g->d_ws + off_vis);
  for (int c = 0; c < 2; ++c) {
    for (int k = 0; k < 3; ++k) g->d_fbm[c][k] = (uint32_t*)(g->d_ws + off_fbm + (size_t)(c * 3 + k) * al(bm_bytes));
    for (int k = 0; k < 2; ++k) {
      g->d_q[c][k] = (int32_t*)(g->d_ws + off_q + (size_t)(c * 2 + k) * al(sizeof(int32_t) * g->qcap));
      g->d_hq[c][k] = (int32_t*)(g->d_ws + off_hq + (size_t)(c * 2 + k) * al(sizeof(int32_t) * g->hcap));
      g->d_hb[c][k] = (int64_t*)(g->d_ws + off_hb + (size_t)(c * 2 + k) * al(sizeof(int64_t) * g->hcap));
    }
  }
  g->d_ctl = (fcd::VisitCtl*)(g->d_ws + off_ctl);
  // canonical parameters, then the copy in the batch kernel's layout
  CK(hipMalloc((void**)&g->d_mlp, 2 * sizeof(float) * fcd::kMlpParams));
  CK(hipMemsetAsync(g->d_mlp, 0, 2 * sizeof(float) * fcd::kMlpParams, g->stream));
  CK(hipHostMalloc((void**)&g->h_phase, 4 * sizeof(int32_t), hipHostMallocDefault));

  int occ = sizeof(OffT) == 8 ? fcd::phase_occupancy<int64_t>() : fcd::phase_occupancy<int32_t>();
  if (occ < 1) return bail(fail(FC_ERR_NOGPU, "phase kernel has no occupancy on this device (missing gfx950 code?)"));
  // 4 resident blocks per CU measured best for BFS and plain s–t at s22
  // (2: 345, 3: 384, 4: 401, 6: 392, 8: 348 GTEPS): more blocks add
  // contention on the launch's shared counters more than they hide latency.
  int per_cu = std::min(occ, 4);
  const char* bpc = std::getenv("FASTCON_BLOCKS_PER_CU");
  if (bpc && std::atoi(bpc) > 0) per_cu = std::min(occ, std::atoi(bpc));
  g->blocks = per_cu * g->cus;
  fc_status s2 = fcd::ensure_queries(g, 1024);
  if (s2 != FC_OK) return bail(s2);
  // the uploads and clears above are ordered on the graph's stream only:
  // finish them before any other stream (a batch lane's) can touch the graph
  CK(hipStreamSynchronize(g->stream));
#undef CK
  *out = g;
  return FC_OK;
}

extern "C" fc_status fc_graph_create_csr32(const int32_t* off, const int32_t* col, int32_t V, int64_t E,
                                           uint32_t flags, fc_graph** out) {
  return graph_create<int32_t>(off, col, V, E, flags, out);
}

extern "C" fc_status fc_graph_create_csr64(const int64_t* off, const int32_t* col, int64_t V, int64_t E,
                                           uint32_t flags, fc_graph** out) {
  return graph_create<int64_t>(off, col, V, E, flags, out);
}

extern "C" fc_status fc_graph_destroy(fc_graph* g) {
  if (!g) return FC_OK;
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  fcd::batch_free(g->batch);
  if (g->gexec) (void)hipGraphExecDestroy(g->gexec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->h_out) (void)hipHostFree(g->h_out);
  if (g->h_phase) (void)hipHostFree(g->h_phase);
  (void)hipFree(g->d_qs);
  (void)hipFree(g->d_qt);
  (void)hipFree(g->d_out);
  (void)hipFree(g->d_mlp);
  (void)hipFree(g->d_ws);
  (void)hipFree(g->d_col);
  (void)hipFree(g->d_off);
  (void)hipFree(g->d_perm);
  (void)hipFree(g->d_iperm);
  (void)hipFree(g->d_lv_out);
  (void)hipFree(g->d_par_out);
  if (g->ev0) (void)hipEventDestroy(g->ev0);
  if (g->ev1) (void)hipEventDestroy(g->ev1);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
  return FC_OK;
}

// Copy device results to the caller on the graph's stream. The graph's
// streams do not wait for the legacy null stream, so no plain hipMemcpy or
// hipMemset touches its memory.
static hipError_t read_back(void* dst, const void* src, size_t bytes, fc_graph* g) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g->stream);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(g->stream);
}

static void add_stats(const fcd::VisitOut& o, fc_stats* s) {
  s->levels += o.levels;
  s->depth += o.depth;
  s->visited_vertices += o.visited_vertices;
  s->visited_edges += o.visited_edges;
  s->expanded_vertices += o.expanded_vertices;
  s->expanded_edges += o.expanded_edges;
  s->topdown_levels += o.topdown;
  s->bottomup_levels += o.bottomup;
  s->max_frontier = std::max(s->max_frontier, o.max_frontier);
  s->launches += o.launches;
}

static fc_status check_out(const fcd::VisitOut& o) {
  if (o.status != 0) return fail(FC_ERR_CAPACITY, "visit kernel: frontier guard tripped (err=%d)", o.err);
  return FC_OK;
}

// Levels (and parents) of a BFS on a relabelled graph back in caller ids:
// out[v] = lv[perm[v]], parents mapped through iperm. One coalesced pass.
// Levels and parents back in caller ids. A BFS writes state and parent only
// for the vertices it reaches; the visited bitmap says which those are.
__global__ void unpermute_kernel(const int32_t* __restrict__ perm, const int32_t* __restrict__ iperm,
                                 const uint32_t* __restrict__ reached, const int32_t* __restrict__ lv,
                                 const int32_t* __restrict__ par, int32_t* __restrict__ lv_out,
                                 int32_t* __restrict__ par_out, int64_t V) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += stride) {
    const int32_t k = perm[v];
    const bool r = (reached[(uint32_t)k >> 5] >> (k & 31)) & 1u;
    lv_out[v] = r ? lv[k] : -1;
    if (par) {
      const int32_t q = r ? par[k] : -1;
      par_out[v] = (uint32_t)q < (uint32_t)V ? iperm[q] : q;
    }
  }
}

extern "C" fc_status fc_bfs(fc_graph* g, int32_t src, int32_t* levels_out, int32_t* parents_out, fc_stats* stats) {
  fc::clear_error();
  if (!g) return fail(FC_ERR_INVALID, "fc_bfs: null graph");
  if (src < 0 || src >= g->V) return fail(FC_ERR_INVALID, "fc_bfs: source %d out of range [0,%lld)", src, (long long)g->V);
  const bool par = parents_out != nullptr;
  const int32_t none = -1;
  const int32_t dsrc = g->relabeled ? g->h_perm[(size_t)src] : src;
  double ms = 0;
  g->levels_valid = g->parents_valid = false;
  fc_status st = fcd::run_visits(g, fcd::kModeBfs, &dsrc, &none, 1, par, &ms);
  if (st != FC_OK) return st;
  const fcd::VisitOut o = g->h_out[0];
  st = check_out(o);
  if (st != FC_OK) return st;
  if (g->relabeled) {
    // the results in caller ids are part of the BFS: timed with it
    const size_t bytes = sizeof(int32_t) * (size_t)g->V;
    if (!g->d_lv_out) FC_HIP(hipMalloc((void**)&g->d_lv_out, bytes));
    if (par && !g->d_par_out) FC_HIP(hipMalloc((void**)&g->d_par_out, bytes));
    FC_HIP(hipEventRecord(g->ev0, g->stream));
    const int blocks = (int)std::min<int64_t>((g->V + 255) / 256, (int64_t)g->cus * 8);
    hipLaunchKernelGGL(unpermute_kernel, dim3(blocks), dim3(256), 0, g->stream, g->d_perm, g->d_iperm, g->d_visited,
                       g->d_state, par ? g->d_parent : nullptr, g->d_lv_out, g->d_par_out, (int64_t)g->V);
    FC_HIP(hipGetLastError());
    FC_HIP(hipEventRecord(g->ev1, g->stream));
    FC_HIP(hipEventSynchronize(g->ev1));
    float fms = 0.f;
    FC_HIP(hipEventElapsedTime(&fms, g->ev0, g->ev1));
    ms += fms;
  }
  g->levels_valid = true;
  g->parents_valid = par;
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->time_ms = ms;
    add_stats(o, stats);
  }
  const int32_t* lv = g->relabeled ? g->d_lv_out : g->d_state;
  const int32_t* pa = g->relabeled ? g->d_par_out : g->d_parent;
  if (levels_out) FC_HIP(read_back(levels_out, lv, sizeof(int32_t) * (size_t)g->V, g));
  if (parents_out) FC_HIP(read_back(parents_out, pa, sizeof(int32_t) * (size_t)g->V, g));
  return FC_OK;
}

extern "C" fc_status fc_graph_levels(fc_graph* g, int32_t* levels_out) {
  fc::clear_error();
  if (!g || !levels_out) return fail(FC_ERR_INVALID, "fc_graph_levels: null argument");
  if (!g->levels_valid) return fail(FC_ERR_INVALID, "fc_graph_levels: no BFS result on the device");
  const int32_t* lv = g->relabeled ? g->d_lv_out : g->d_state;
  FC_HIP(read_back(levels_out, lv, sizeof(int32_t) * (size_t)g->V, g));
  return FC_OK;
}

extern "C" fc_status fc_graph_parents(fc_graph* g, int32_t* parents_out) {
  fc::clear_error();
  if (!g || !parents_out) return fail(FC_ERR_INVALID, "fc_graph_parents: null argument");
  if (!g->parents_valid) return fail(FC_ERR_INVALID, "fc_graph_parents: the last BFS did not record parents");
  const int32_t* pa = g->relabeled ? g->d_par_out : g->d_parent;
  FC_HIP(read_back(parents_out, pa, sizeof(int32_t) * (size_t)g->V, g));
  return FC_OK;
}

static fc_status stcon_run(fc_graph* g, const int32_t* s, const int32_t* t, int32_t n, int32_t* connected,
                           int32_t* dist, int32_t* meet, fc_stats* stats) {
  fc::clear_error();
  if (!g || (n > 0 && (!s || !t))) return fail(FC_ERR_INVALID, "s–t query: null argument");
  if (n < 0) return fail(FC_ERR_INVALID, "s–t query: negative count");
  for (int32_t i = 0; i < n; ++i)
    if (s[i] < 0 || s[i] >= g->V || t[i] < 0 || t[i] >= g->V)
      return fail(FC_ERR_INVALID, "s–t query %d: vertex out of range [0,%lld)", i, (long long)g->V);
  g->levels_valid = g->parents_valid = false;
  // s == t needs no visit: distance 0.
  std::vector<int32_t> vs, vt, idx;
  vs.reserve(n);
  vt.reserve(n);
  idx.reserve(n);
  for (int32_t i = 0; i < n; ++i)
    if (s[i] != t[i]) {
      vs.push_back(g->relabeled ? g->h_perm[(size_t)s[i]] : s[i]);
      vt.push_back(g->relabeled ? g->h_perm[(size_t)t[i]] : t[i]);
      idx.push_back(i);
    }
  double ms = 0;
  fc_status st = fcd::run_visits(g, fcd::kModePlainSt, vs.data(), vt.data(), (int32_t)vs.size(), false, &ms);
  if (st != FC_OK) return st;
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->time_ms = ms;
  }
  for (int32_t i = 0; i < n; ++i) {
    if (connected) connected[i] = 1;
    if (dist) dist[i] = 0;
    if (meet) meet[i] = s[i];
  }
  for (size_t k = 0; k < idx.size(); ++k) {
    const fcd::VisitOut o = g->h_out[k];
    st = check_out(o);
    if (st != FC_OK) return st;
    const int32_t i = idx[k];
    if (connected) connected[i] = o.dist >= 0 ? 1 : 0;
    if (dist) dist[i] = o.dist;
    if (meet) meet[i] = (g->relabeled && (uint32_t)o.meet < (uint32_t)g->V) ? g->h_iperm[(size_t)o.meet] : o.meet;
    if (stats) add_stats(o, stats);
  }
  return FC_OK;
}

extern "C" fc_status fc_stcon(fc_graph* g, int32_t s, int32_t t, int32_t* connected, int32_t* dist, int32_t* meet,
                              fc_stats* stats) {
  return stcon_run(g, &s, &t, 1, connected, dist, meet, stats);
}

extern "C" fc_status fc_stcon_many(fc_graph* g, const int32_t* s, const int32_t* t, int32_t n, int32_t* connected,
                                   int32_t* dist, fc_stats* stats) {
  return stcon_run(g, s, t, n, connected, dist, nullptr, stats);
}

// Shortest s–t path from the parents a plain visit recorded (one wave).
// The meeting vertex u has level lu in its own colour's tree; the other end
// of the meeting edge is a neighbour w of u of the other colour at level
// dist - 1 - lu. The lanes search u's adjacency for w; lane 0 then walks
// both parent chains (a root is its own parent) and writes the path, s
// first, into out[0..dist]. *status receives dist + 1, or a negative code:
// -1 no such neighbour, -2 a chain left the vertex range, -3 a chain did not
// end at a root.
template <typename OffT>
__global__ void path_kernel(const OffT* __restrict__ off, const int32_t* __restrict__ col,
                            const unsigned long long* __restrict__ sst, uint32_t epoch,
                            const int32_t* __restrict__ parent, int32_t V, int32_t u, int32_t dist,
                            int32_t* __restrict__ out, int32_t* __restrict__ status) {
  using namespace fcd;
  const int lane = lane_id();
  const unsigned long long wu = sst[u];
  const int32_t su = (int32_t)(uint32_t)wu;
  const bool tu = (su & kColorBit) != 0;  // u on the t side
  const int32_t lu = su & kLevelMask;
  const int32_t want = dist - 1 - lu;
  int32_t w = -1;
  const OffT e1 = off[u + 1];
  for (OffT b = off[u]; b < e1; b += kWave) {
    const OffT e = b + (OffT)lane;
    int32_t x = -1;
    bool hit = false;
    if (e < e1) {
      x = col[e];
      const unsigned long long wx = sst[x];
      const int32_t sx = (int32_t)(uint32_t)wx;
      hit = (uint32_t)(wx >> 32) == epoch && (((sx & kColorBit) != 0) != tu) && (sx & kLevelMask) == want;
    }
    const uint64_t m = ballot(hit);
    if (m) {
      w = __shfl(x, __ffsll((unsigned long long)m) - 1);
      break;
    }
  }
  if (lane != 0) return;
  if (w < 0 || want < 0 || (uint32_t)(wu >> 32) != epoch) {
    *status = -1;
    return;
  }
  const int32_t a = tu ? w : u, b = tu ? u : w;  // s-side and t-side ends
  const int32_t la = tu ? want : lu;             // level of a from s
  int32_t x = a;
  for (int32_t k = la; k >= 0; --k) {
    if ((uint32_t)x >= (uint32_t)V) {
      *status = -2;
      return;
    }
    out[k] = x;
    x = parent[x];
  }
  if (x != out[0]) {  // out[0] must be the s root, its own parent
    *status = -3;
    return;
  }
  x = b;
  for (int32_t k = la + 1; k <= dist; ++k) {
    if ((uint32_t)x >= (uint32_t)V) {
      *status = -2;
      return;
    }
    out[k] = x;
    x = parent[x];
  }
  if (x != out[dist]) {  // out[dist] must be the t root
    *status = -3;
    return;
  }
  *status = dist + 1;
}

extern "C" fc_status fc_stcon_path(fc_graph* g, int32_t s, int32_t t, int32_t* path_out, int32_t cap, int32_t* len_out,
                                   fc_stats* stats) {
  fc::clear_error();
  if (!g || !len_out || (cap > 0 && !path_out)) return fail(FC_ERR_INVALID, "fc_stcon_path: null argument");
  if (s < 0 || s >= g->V || t < 0 || t >= g->V)
    return fail(FC_ERR_INVALID, "fc_stcon_path: vertex out of range [0,%lld)", (long long)g->V);
  g->levels_valid = g->parents_valid = false;
  if (stats) std::memset(stats, 0, sizeof(*stats));
  *len_out = 0;
  if (s == t) {
    *len_out = 1;
    if (cap < 1) return fail(FC_ERR_CAPACITY, "fc_stcon_path: the path needs 1 entry, cap is %d", cap);
    path_out[0] = s;
    return FC_OK;
  }
  const int32_t ds = g->relabeled ? g->h_perm[(size_t)s] : s;
  const int32_t dt = g->relabeled ? g->h_perm[(size_t)t] : t;
  double ms = 0;
  fc_status st = fcd::run_visits(g, fcd::kModePlainSt, &ds, &dt, 1, true, &ms);
  if (st != FC_OK) return st;
  const uint32_t epoch = g->epoch_next - 1;  // the query's tag in the state words
  const fcd::VisitOut o = g->h_out[0];
  st = check_out(o);
  if (st != FC_OK) return st;
  if (stats) {
    stats->time_ms = ms;
    add_stats(o, stats);
  }
  if (o.dist < 0) return FC_OK;
  *len_out = o.dist + 1;
  if (o.dist + 1 > cap)
    return fail(FC_ERR_CAPACITY, "fc_stcon_path: the path needs %d entries, cap is %d", o.dist + 1, cap);
  if ((uint32_t)o.meet >= (uint32_t)g->V || o.dist + 1 > g->V)
    return fail(FC_ERR_CAPACITY, "fc_stcon_path: visit returned meet %d at distance %d", o.meet, o.dist);
  // The frontier queues are idle after the visit: the path goes to the
  // first, the status word to the second.
  int32_t* d_path = g->d_q[0][0];
  int32_t* d_status = g->d_q[1][0];
  FC_HIP(hipEventRecord(g->ev0, g->stream));
  if (g->off64)
    hipLaunchKernelGGL(path_kernel<int64_t>, dim3(1), dim3(fcd::kWave), 0, g->stream, (const int64_t*)g->d_off,
                       g->d_col, g->d_sst, epoch, g->d_parent, (int32_t)g->V, o.meet, o.dist, d_path, d_status);
  else
    hipLaunchKernelGGL(path_kernel<int32_t>, dim3(1), dim3(fcd::kWave), 0, g->stream, (const int32_t*)g->d_off,
                       g->d_col, g->d_sst, epoch, g->d_parent, (int32_t)g->V, o.meet, o.dist, d_path, d_status);
  FC_HIP(hipGetLastError());
  FC_HIP(hipEventRecord(g->ev1, g->stream));
  int32_t status = 0;
  FC_HIP(hipMemcpyAsync(&status, d_status, sizeof(int32_t), hipMemcpyDeviceToHost, g->stream));
  FC_HIP(hipMemcpyAsync(path_out, d_path, sizeof(int32_t) * (size_t)(o.dist + 1), hipMemcpyDeviceToHost, g->stream));
  FC_HIP(hipStreamSynchronize(g->stream));
  float pms = 0.f;
  FC_HIP(hipEventElapsedTime(&pms, g->ev0, g->ev1));
  if (stats) stats->time_ms += pms;
  if (status != o.dist + 1) return fail(FC_ERR_HIP, "fc_stcon_path: path assembly failed (code %d)", status);
  if (g->relabeled)
    for (int32_t k = 0; k <= o.dist; ++k) path_out[k] = g->h_iperm[(size_t)path_out[k]];
  if (path_out[0] != s || path_out[o.dist] != t)
    return fail(FC_ERR_HIP, "fc_stcon_path: assembled path runs %d..%d", path_out[0], path_out[o.dist]);
  return FC_OK;
}

// validate_stcon_bfs (st-conn/make_stcon.c:1578-1641) on one GPU: the BFS
// from mn runs on the device, the parent walks on the host.
extern "C" fc_status fc_validate_stcon(fc_graph* g, int32_t s, int32_t t, int32_t mn, int32_t depth) {
  fc::clear_error();
  if (!g) return fail(FC_ERR_INVALID, "fc_validate_stcon: null graph");
  if (s < 0 || s >= g->V || t < 0 || t >= g->V || mn < 0 || mn >= g->V)
    return fail(FC_ERR_INVALID, "fc_validate_stcon: vertex out of range [0,%lld)", (long long)g->V);
  std::vector<int32_t> lv((size_t)g->V), par((size_t)g->V);
  fc_status st = fc_bfs(g, mn, lv.data(), par.data(), nullptr);
  if (st != FC_OK) return st;
  int32_t dep[2] = {-1, -1};
  const int32_t from[2] = {t, s};  // blue path first, as the reference
  for (int k = 0; k < 2; ++k) {
    int32_t v = from[k], p = v;
    int64_t steps = -1;
    do {
      v = p;
      if (v < 0 || v >= g->V) return fail(FC_ERR_MISMATCH, "%s path leaves the graph", k ? "RED" : "BLUE");
      p = par[(size_t)v];
      ++steps;
    } while (v != p && steps <= g->V);
    if (v != mn)
      return fail(FC_ERR_MISMATCH, "WRONG MATCHING NODE: %s Path leaf = %d, Matching Node = %d", k ? "RED" : "BLUE", v,
                  mn);
    dep[k] = (int32_t)steps;
    if (steps > depth)
      return fail(FC_ERR_MISMATCH, "WRONG ST-CON depth on %s: ST-CON depth = %d, bfs depth = %d", k ? "RED" : "BLUE",
                  depth, dep[k]);
  }
  const int32_t mx = std::max(dep[0], dep[1]);
  if (mx != depth)
    return fail(FC_ERR_MISMATCH, "WRONG ST-CON MAX depth: ST-CON MAX depth = %d, bfs MAX depth = %d", depth, mx);
  return FC_OK;
}

extern "C" fc_status fc_stcon_sched_many(fc_graph* g, const int32_t* s, const int32_t* t, int32_t n, int32_t policy,
                                         int32_t* connected, int32_t* dist, fc_stats* stats) {
  if (policy == FC_POLICY_PLAIN) return fc_stcon_many(g, s, t, n, connected, dist, stats);
  if (policy != FC_POLICY_HEURISTIC && policy != FC_POLICY_LEARNED)
    return fail(FC_ERR_INVALID, "unknown policy %d", policy);
  // One-sided scheduled steps for single queries are not built: an earlier
  // persistent-kernel version of this path waited out its grid-barrier
  // bound on the GPU, that cause was not identified, and the code was
  // removed. Heuristic and learned scheduling run in the batched engine.
  return fail(FC_ERR_UNSUPPORTED,
              "scheduled single-query visits are switched off; use fc_stcon_batch for heuristic or "
              "learned scheduling");
}
