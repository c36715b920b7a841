// C ABI of the device engine: graph upload, BFS and plain s–t queries.
//
// fc_graph_create_* replaces cudaGraph::cudaGraph (BFS-4K/src/device/cudaGraph.cu:7-36):
// the CSR is uploaded once and every per-query buffer is allocated up front,
// sized for the graph (queues of V entries per colour and parity, three
// frontier bitmaps per colour, a hub queue sized by the number of vertices of
// degree >= hub_threshold) instead of the reference's free/2 frontier buffers.
// fc_bfs replaces Reset + cudaBFS4K (BFS-4K/src/device/BFS_WorkEfficient.cu:81-133);
// fc_stcon replaces make_stcon's single-GPU path (st-conn/make_stcon.c:730-1497).
#include "engine.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>

using fc::fail;

namespace fc {
fc_status validate_csr32(const int32_t* off, const int32_t* col, int64_t V, int64_t E, int64_t hub, int64_t* hubs);
fc_status validate_csr64(const int64_t* off, const int32_t* col, int64_t V, int64_t E, int64_t hub, int64_t* hubs);
}  // namespace fc

namespace fcd {

fc_status hip_fail(hipError_t e, const char* what) {
  return fail(e == hipErrorNoDevice || e == hipErrorNoBinaryForGpu || e == hipErrorInvalidDevice ? FC_ERR_NOGPU
                                                                                                  : FC_ERR_HIP,
              "%s: %s", what, hipGetErrorString(e));
}

fc_status ensure_out(fc_graph* g, int32_t n) {
  if (n <= g->out_cap) return FC_OK;
  if (g->h_out) FC_HIP(hipHostFree(g->h_out));
  g->h_out = nullptr;
  g->out_cap = 0;
  int32_t cap = std::max<int32_t>(n, 64);
  FC_HIP(hipHostMalloc((void**)&g->h_out, sizeof(VisitOut) * (size_t)cap, hipHostMallocDefault));
  g->out_cap = cap;
  return FC_OK;
}

template <typename OffT>
static VisitParams<OffT> make_params(fc_graph* g) {
  VisitParams<OffT> p;
  std::memset(&p, 0, sizeof(p));
  p.off = (const OffT*)g->d_off;
  p.col = g->d_col;
  p.V = (int32_t)g->V;
  p.E = g->E;
  p.hub_threshold = g->hub_threshold;
  p.state = g->d_state;
  p.parent = nullptr;
  p.visited = g->d_visited;
  for (int c = 0; c < 2; ++c) {
    for (int k = 0; k < 3; ++k) p.fbm[c][k] = g->d_fbm[c][k];
    for (int k = 0; k < 2; ++k) {
      p.q[c][k] = g->d_q[c][k];
      p.hq[c][k] = g->d_hq[c][k];
      p.hb[c][k] = g->d_hb[c][k];
    }
  }
  p.qcap = g->qcap;
  p.hcap = g->hcap;
  p.ctl = g->d_ctl;
  p.spin_limit = g->spin_limit;
  p.force_side = -1;
  p.force_strategy = -1;
  p.trace_slot = -1;
  return p;
}

fc_status launch_one(fc_graph* g, int mode, int policy, int32_t src, int32_t dst, bool parents, int slot,
                     int force_side, int force_strategy, int max_steps, bool resume, int lvl0, int lvl1,
                     float* trace_feat) {
  hipError_t e;
  if (g->off64) {
    VisitParams<int64_t> p = make_params<int64_t>(g);
    p.parent = parents ? g->d_parent : nullptr;
    p.out = g->h_out + slot;
    p.src = src;
    p.dst = dst;
    p.mode = mode;
    p.policy = policy;
    p.mlp = (policy == 2 && g->mlp_set) ? g->d_mlp : nullptr;
    p.force_side = force_side;
    p.force_strategy = force_strategy;
    p.max_steps = max_steps;
    p.resume = resume ? 1 : 0;
    p.lvl0 = lvl0;
    p.lvl1 = lvl1;
    p.trace_feat = trace_feat;
    e = launch_visit<int64_t>(p, g->blocks, g->stream);
  } else {
    VisitParams<int32_t> p = make_params<int32_t>(g);
    p.parent = parents ? g->d_parent : nullptr;
    p.out = g->h_out + slot;
    p.src = src;
    p.dst = dst;
    p.mode = mode;
    p.policy = policy;
    p.mlp = (policy == 2 && g->mlp_set) ? g->d_mlp : nullptr;
    p.force_side = force_side;
    p.force_strategy = force_strategy;
    p.max_steps = max_steps;
    p.resume = resume ? 1 : 0;
    p.lvl0 = lvl0;
    p.lvl1 = lvl1;
    p.trace_feat = trace_feat;
    e = launch_visit<int32_t>(p, g->blocks, g->stream);
  }
  if (e != hipSuccess) return hip_fail(e, "visit kernel launch");
  return FC_OK;
}

}  // namespace fcd

using fcd::ensure_out;
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
  const char* ht = std::getenv("FASTCON_HUB_THRESHOLD");
  if (ht && std::atoi(ht) > 0) g->hub_threshold = std::atoi(ht);
  int64_t hubs = 0;
  fc_status st = sizeof(OffT) == 8
                     ? fc::validate_csr64((const int64_t*)off, col, V, E, g->hub_threshold, &hubs)
                     : fc::validate_csr32((const int32_t*)off, col, V, E, g->hub_threshold, &hubs);
  if (st != FC_OK) {
    delete g;
    return st;
  }
  auto bail = [&](fc_status s) {
    fc_graph_destroy(g);
    return s;
  };
#define CK(call)                                  \
  do {                                            \
    hipError_t _e = (call);                       \
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
  CK(hipMemcpy(g->d_off, off, sizeof(OffT) * (size_t)(V + 1), hipMemcpyHostToDevice));
  if (E > 0) CK(hipMemcpy(g->d_col, col, sizeof(int32_t) * (size_t)E, hipMemcpyHostToDevice));

  // Workspace layout (256-byte aligned pieces of one allocation).
  const size_t nw64 = (size_t)((V + 63) / 64);
  const size_t bm_bytes = nw64 * 2 * sizeof(uint32_t);
  g->qcap = (uint32_t)V;
  g->hcap = (uint32_t)std::max<int64_t>(hubs, 1);
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  size_t off_state = 0;
  size_t off_parent = off_state + al(sizeof(int32_t) * (size_t)V);
  size_t off_vis = off_parent + al(sizeof(int32_t) * (size_t)V);
  size_t off_fbm = off_vis + al(bm_bytes);
  size_t off_q = off_fbm + 6 * al(bm_bytes);
  size_t off_hq = off_q + 4 * al(sizeof(int32_t) * (size_t)g->qcap);
  size_t off_hb = off_hq + 4 * al(sizeof(int32_t) * (size_t)g->hcap);
  size_t off_ctl = off_hb + 4 * al(sizeof(int64_t) * (size_t)g->hcap);
  g->ws_bytes = off_ctl + al(sizeof(fcd::VisitCtl));
  CK(hipMalloc((void**)&g->d_ws, g->ws_bytes));
  CK(hipMemset(g->d_ws, 0, g->ws_bytes));
  g->d_state = (int32_t*)(g->d_ws + off_state);
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
  CK(hipMalloc((void**)&g->d_mlp, sizeof(float) * fcd::kMlpParams));

  int occ = sizeof(OffT) == 8 ? fcd::visit_occupancy<int64_t>() : fcd::visit_occupancy<int32_t>();
  if (occ < 1) return bail(fail(FC_ERR_NOGPU, "visit kernel has no occupancy on this device (missing gfx950 code?)"));
  int per_cu = std::min(occ, 4);
  const char* bpc = std::getenv("FASTCON_BLOCKS_PER_CU");
  if (bpc && std::atoi(bpc) > 0) per_cu = std::min(occ, std::atoi(bpc));
  g->blocks = per_cu * g->cus;
  // Barrier waits give up after ~10 s of real time (100 MHz counter).
  g->spin_limit = 1000000000ull;
  const char* sl = std::getenv("FASTCON_SPIN_LIMIT");
  if (sl && std::atoll(sl) > 0) g->spin_limit = (uint64_t)std::atoll(sl);
  fc_status s2 = ensure_out(g, 64);
  if (s2 != FC_OK) return bail(s2);
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
  if (g->h_out) (void)hipHostFree(g->h_out);
  (void)hipFree(g->d_mlp);
  (void)hipFree(g->d_snap);
  (void)hipFree(g->d_ws);
  (void)hipFree(g->d_col);
  (void)hipFree(g->d_off);
  if (g->ev0) (void)hipEventDestroy(g->ev0);
  if (g->ev1) (void)hipEventDestroy(g->ev1);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
  return FC_OK;
}

static void fill_stats(const fcd::VisitOut& o, double ms, fc_stats* s) {
  if (!s) return;
  std::memset(s, 0, sizeof(*s));
  s->time_ms = ms;
  s->levels = o.levels;
  s->depth = o.depth;
  s->visited_vertices = o.visited_vertices;
  s->visited_edges = o.visited_edges;
  s->expanded_vertices = o.expanded_vertices;
  s->expanded_edges = o.expanded_edges;
  s->topdown_levels = o.topdown;
  s->bottomup_levels = o.bottomup;
  s->max_frontier = o.max_frontier;
}

static fc_status check_out(fc_graph* g, const fcd::VisitOut& o) {
  if (o.status == fcd::FC_TIMEOUT_STATUS) {
    // Leave the barrier usable for the next launch.
    (void)hipMemsetAsync(g->d_ctl, 0, sizeof(fcd::GridBar), g->stream);
    (void)hipStreamSynchronize(g->stream);
    return fail(FC_ERR_TIMEOUT, "visit kernel: grid barrier wait exceeded its bound; results are invalid");
  }
  if (o.status != 0) return fail(FC_ERR_CAPACITY, "visit kernel: frontier guard tripped (err=%d)", o.err);
  return FC_OK;
}

extern "C" fc_status fc_bfs(fc_graph* g, int32_t src, int32_t* levels_out, int32_t* parents_out, fc_stats* stats) {
  fc::clear_error();
  if (!g) return fail(FC_ERR_INVALID, "fc_bfs: null graph");
  if (src < 0 || src >= g->V) return fail(FC_ERR_INVALID, "fc_bfs: source %d out of range [0,%lld)", src, (long long)g->V);
  const bool par = parents_out != nullptr;
  FC_HIP(hipEventRecord(g->ev0, g->stream));
  fc_status st = fcd::launch_one(g, fcd::kModeBfs, 0, src, -1, par, 0);
  if (st != FC_OK) return st;
  FC_HIP(hipEventRecord(g->ev1, g->stream));
  FC_HIP(hipEventSynchronize(g->ev1));
  float ms = 0.f;
  FC_HIP(hipEventElapsedTime(&ms, g->ev0, g->ev1));
  const fcd::VisitOut o = g->h_out[0];
  st = check_out(g, o);
  g->levels_valid = st == FC_OK;
  g->parents_valid = st == FC_OK && par;
  if (st != FC_OK) return st;
  fill_stats(o, ms, stats);
  if (levels_out) FC_HIP(hipMemcpy(levels_out, g->d_state, sizeof(int32_t) * (size_t)g->V, hipMemcpyDeviceToHost));
  if (parents_out) FC_HIP(hipMemcpy(parents_out, g->d_parent, sizeof(int32_t) * (size_t)g->V, hipMemcpyDeviceToHost));
  return FC_OK;
}

extern "C" fc_status fc_graph_levels(fc_graph* g, int32_t* levels_out) {
  fc::clear_error();
  if (!g || !levels_out) return fail(FC_ERR_INVALID, "fc_graph_levels: null argument");
  if (!g->levels_valid) return fail(FC_ERR_INVALID, "fc_graph_levels: no BFS result on the device");
  FC_HIP(hipMemcpy(levels_out, g->d_state, sizeof(int32_t) * (size_t)g->V, hipMemcpyDeviceToHost));
  return FC_OK;
}

extern "C" fc_status fc_graph_parents(fc_graph* g, int32_t* parents_out) {
  fc::clear_error();
  if (!g || !parents_out) return fail(FC_ERR_INVALID, "fc_graph_parents: null argument");
  if (!g->parents_valid) return fail(FC_ERR_INVALID, "fc_graph_parents: the last BFS did not record parents");
  FC_HIP(hipMemcpy(parents_out, g->d_parent, sizeof(int32_t) * (size_t)g->V, hipMemcpyDeviceToHost));
  return FC_OK;
}

static fc_status stcon_run(fc_graph* g, const int32_t* s, const int32_t* t, int32_t n, int mode, int policy,
                           int32_t* connected, int32_t* dist, int32_t* meet, fc_stats* stats) {
  fc::clear_error();
  if (!g || (n > 0 && (!s || !t))) return fail(FC_ERR_INVALID, "s–t query: null argument");
  if (n < 0) return fail(FC_ERR_INVALID, "s–t query: negative count");
  for (int32_t i = 0; i < n; ++i)
    if (s[i] < 0 || s[i] >= g->V || t[i] < 0 || t[i] >= g->V)
      return fail(FC_ERR_INVALID, "s–t query %d: vertex out of range [0,%lld)", i, (long long)g->V);
  fc_status st = ensure_out(g, n);
  if (st != FC_OK) return st;
  g->levels_valid = g->parents_valid = false;
  FC_HIP(hipEventRecord(g->ev0, g->stream));
  for (int32_t i = 0; i < n; ++i) {
    if (s[i] == t[i]) continue;
    st = fcd::launch_one(g, mode, policy, s[i], t[i], false, i);
    if (st != FC_OK) return st;
  }
  FC_HIP(hipEventRecord(g->ev1, g->stream));
  FC_HIP(hipEventSynchronize(g->ev1));
  float ms = 0.f;
  FC_HIP(hipEventElapsedTime(&ms, g->ev0, g->ev1));
  fcd::VisitOut agg;
  std::memset(&agg, 0, sizeof(agg));
  for (int32_t i = 0; i < n; ++i) {
    fcd::VisitOut o;
    if (s[i] == t[i]) {
      std::memset(&o, 0, sizeof(o));
      o.dist = 0;
      o.meet = s[i];
      o.visited_vertices = 1;
    } else {
      o = g->h_out[i];
      st = check_out(g, o);
      if (st != FC_OK) return st;
    }
    if (connected) connected[i] = o.dist >= 0 ? 1 : 0;
    if (dist) dist[i] = o.dist;
    if (meet) meet[i] = o.meet;
    agg.levels += o.levels;
    agg.depth += o.depth;
    agg.visited_vertices += o.visited_vertices;
    agg.visited_edges += o.visited_edges;
    agg.expanded_vertices += o.expanded_vertices;
    agg.expanded_edges += o.expanded_edges;
    agg.topdown += o.topdown;
    agg.bottomup += o.bottomup;
    agg.max_frontier = std::max(agg.max_frontier, o.max_frontier);
  }
  fill_stats(agg, ms, stats);
  return FC_OK;
}

extern "C" fc_status fc_stcon(fc_graph* g, int32_t s, int32_t t, int32_t* connected, int32_t* dist, int32_t* meet,
                              fc_stats* stats) {
  return stcon_run(g, &s, &t, 1, fcd::kModePlainSt, 0, connected, dist, meet, stats);
}

extern "C" fc_status fc_stcon_many(fc_graph* g, const int32_t* s, const int32_t* t, int32_t n, int32_t* connected,
                                   int32_t* dist, fc_stats* stats) {
  return stcon_run(g, s, t, n, fcd::kModePlainSt, 0, connected, dist, nullptr, stats);
}

extern "C" fc_status fc_stcon_sched_many(fc_graph* g, const int32_t* s, const int32_t* t, int32_t n, int32_t policy,
                                         int32_t* connected, int32_t* dist, fc_stats* stats) {
  if (policy == FC_POLICY_PLAIN) return fc_stcon_many(g, s, t, n, connected, dist, stats);
  if (policy != FC_POLICY_HEURISTIC && policy != FC_POLICY_LEARNED)
    return fail(FC_ERR_INVALID, "unknown policy %d", policy);
  // One-sided scheduled steps inside the persistent kernel are switched off:
  // on the GPU such a launch waited out its grid-barrier bound and the cause
  // has not been identified. Scheduled visits run in the batched engine.
  return fail(FC_ERR_UNSUPPORTED,
              "scheduled single-query visits are disabled (grid-barrier timeout under investigation); "
              "use fc_stcon_batch for heuristic or learned scheduling");
}
