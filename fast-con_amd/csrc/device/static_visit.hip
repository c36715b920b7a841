// fc_stcon_many_static: plain s–t queries on a CSR graph that already sits in
// the caller's device memory, with every per-query buffer taken from one
// fixed device array, so that a call allocates nothing: no hipMalloc, no
// hipFree, no stream, event or graph creation. It is for callers that must
// not allocate device memory (allocation-free serving loops, harnesses that
// own all device memory themselves). The reference has no such entry point:
// its ST-CON wrappers cudaMalloc and cudaFree their scratch on every call
// (st-conn/make_stcon_func.cu:59-107).
//
// The visit is the phase machine of visit.hip in its plain s–t mode (two
// colours, top-down, one atomicCAS per claimed vertex), on the caller's
// numbering and adjacency order; verdicts and distances equal fc_stcon_many's.
//
// The array `scratch` below is a file-scope __device__ object: every GPU the
// library runs on holds its own copy, at its own address, for the life of
// the process. Host code never uses the name as an address. The first call
// on a device looks the address up with hipGetSymbolAddress (with that
// device current) and keeps it; kernels get that address as an argument.
// Every call on one GPU shares that GPU's array, so calls that overlap on one
// GPU (from several host threads) race: serialise them. Each call writes the
// control block and the query list before its first launch, and every query
// tags the state words it writes with an epoch of its own (a word of another
// epoch counts as unvisited), so nothing left by an earlier call (a larger
// graph, more queries, an aborted run) is read as this call's.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <mutex>

#include "engine.hpp"

using fc::fail;

namespace fcd {
namespace {

// Limits. Hubs (degree >= kStaticHub) number at most E / kStaticHub, so with
// E <= kStaticMaxE they always fit the kStaticHubCap-entry hub queues.
constexpr int64_t kStaticMaxV = 1ll << 20;
constexpr int64_t kStaticMaxE = 1ll << 25;
constexpr int32_t kStaticHub = 512;
constexpr int64_t kStaticHubCap = kStaticMaxE / kStaticHub;
constexpr int32_t kStaticQueries = 1024;  // queries per pass of the phase machine
constexpr int kMaxDevices = 64;

constexpr size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }
// Layout of the array: state words, four frontier queues, four hub queues
// (vertex and edge base), the control block, the query lists and the results.
constexpr size_t kOffSst = 0;
constexpr size_t kOffQ = kOffSst + al256(sizeof(unsigned long long) * kStaticMaxV);
constexpr size_t kOffHq = kOffQ + 4 * al256(sizeof(int32_t) * kStaticMaxV);
constexpr size_t kOffHb = kOffHq + 4 * al256(sizeof(int32_t) * kStaticHubCap);
constexpr size_t kOffCtl = kOffHb + 4 * al256(sizeof(int64_t) * kStaticHubCap);
constexpr size_t kOffQs = kOffCtl + al256(sizeof(VisitCtl));
constexpr size_t kOffQt = kOffQs + al256(sizeof(int32_t) * kStaticQueries);
constexpr size_t kOffOut = kOffQt + al256(sizeof(int32_t) * kStaticQueries);
constexpr size_t kStaticBytes = kOffOut + al256(sizeof(VisitOut) * kStaticQueries);

}  // namespace

// The workspace of fc_stcon_many_static (about 27 MB per device).
static __device__ __align__(256) unsigned char scratch[kStaticBytes];

namespace {

struct DeviceSlot {
  unsigned char* ptr = nullptr;  // device address of `scratch` on this device
  int blocks = 0;
  double est_launches = 3.0;     // working launches per query (running estimate)
  uint32_t epoch_next = 1;       // epoch of the next query (0 = never visited)
};
DeviceSlot g_slots[kMaxDevices];
std::mutex g_slots_mu;

// Address of this device's array and the visit grid, looked up once per
// device with that device current.
fc_status device_slot(DeviceSlot** out) {
  int dev = -1;
  FC_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(FC_ERR_NOGPU, "fc_stcon_many_static: device %d not supported", dev);
  std::lock_guard<std::mutex> lk(g_slots_mu);
  DeviceSlot& s = g_slots[dev];
  if (!s.ptr) {
    void* p = nullptr;
    FC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(FC_ERR_NOGPU, "fc_stcon_many_static: no device address for the workspace array");
    int cus = 0;
    FC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int occ = phase_occupancy<int32_t>();
    if (occ < 1 || cus < 1) return fail(FC_ERR_NOGPU, "phase kernel has no occupancy on this device (missing gfx950 code?)");
    s.blocks = std::min(occ, 4) * cus;  // the grid fc_graph_create uses
    s.ptr = (unsigned char*)p;
  }
  *out = &s;
  return FC_OK;
}

VisitParams<int32_t> static_params(unsigned char* base, const int32_t* off, const int32_t* col, int32_t V,
                                   int64_t E) {
  VisitParams<int32_t> p;
  std::memset(&p, 0, sizeof(p));
  set_param(p.off, off);
  set_param(p.col, col);
  p.V = V;
  p.hub_threshold = kStaticHub;
  p.E = E;
  set_param(p.sst, (unsigned long long*)(base + kOffSst));
  // The plain s–t mode keeps no levels, parents or bitmaps (visit.hip):
  // those pointers stay null and are never dereferenced in this mode.
  for (int c = 0; c < 2; ++c)
    for (int k = 0; k < 2; ++k) {
      const int i = c * 2 + k;
      set_param(p.q[c][k], (int32_t*)(base + kOffQ + (size_t)i * al256(sizeof(int32_t) * kStaticMaxV)));
      set_param(p.hq[c][k], (int32_t*)(base + kOffHq + (size_t)i * al256(sizeof(int32_t) * kStaticHubCap)));
      set_param(p.hb[c][k], (int64_t*)(base + kOffHb + (size_t)i * al256(sizeof(int64_t) * kStaticHubCap)));
    }
  p.qcap = (uint32_t)V;
  p.hcap = (uint32_t)kStaticHubCap;
  set_param(p.ctl, (VisitCtl*)(base + kOffCtl));
  set_param(p.qs, (const int32_t*)(base + kOffQs));
  set_param(p.qt, (const int32_t*)(base + kOffQt));
  set_param(p.out, (VisitOut*)(base + kOffOut));
  p.lazy_state = 0;
  p.v_active = V;
  return p;
}

// One pass of the phase machine over n <= kStaticQueries queries already in
// the array; the results are left in p.out. Launches go to the null stream
// (parity 0, 1, 0, ...); the phase word is polled with synchronous copies
// into host stack memory.
fc_status static_pass(DeviceSlot& slot, const VisitParams<int32_t>& p, int32_t n, const int32_t* s, const int32_t* t) {
  if ((uint64_t)slot.epoch_next + (uint64_t)n >= 0xFFFFFFF0ull) {
    // the epoch counter would wrap: clear the state words first
    FC_HIP(hipMemset(p.sst, 0, sizeof(unsigned long long) * kStaticMaxV));
    slot.epoch_next = 1;
  }
  const VisitCtl init = visit_init(p, kModePlainSt, n, s, t, slot.epoch_next, 0, false);
  slot.epoch_next += (uint32_t)n;
  FC_HIP(hipMemcpy(p.ctl, &init, sizeof(init), hipMemcpyHostToDevice));
  int64_t expected = (int64_t)((double)n * slot.est_launches) + 2;
  int64_t launched = 0;
  for (int round = 0;; ++round) {
    int64_t want = std::min<int64_t>(std::max<int64_t>(expected - launched, 1), kPollLaunches);
    for (int64_t i = 0; i < want; ++i, ++launched) {
      hipError_t e = launch_phase<int32_t>(p.ctl, (int)(launched & 1), slot.blocks, nullptr);
      if (e != hipSuccess) return hip_fail(e, "phase kernel launch");
    }
    int32_t word[2] = {0, 0};
    FC_HIP(hipMemcpy(word, &p.ctl->st[launched & 1], sizeof(word), hipMemcpyDeviceToHost));
    if (word[0] == kPhaseIdle) break;
    if (launched > (int64_t)n * ((int64_t)p.V + 4) + 64)
      return fail(FC_ERR_CAPACITY, "phase machine did not finish after %lld launches", (long long)launched);
    if (launched < expected) continue;
    const int64_t remaining = std::max<int64_t>(n - word[1], 1);
    expected = launched + (int64_t)((double)remaining * (slot.est_launches + 1.0) * (1.0 + round)) + 1;
  }
  return FC_OK;
}

}  // namespace
}  // namespace fcd

extern "C" fc_status fc_stcon_many_static(const int32_t* d_off, const int32_t* d_col, int32_t V, int64_t E,
                                          const int32_t* s, const int32_t* t, int32_t n, int32_t* connected,
                                          int32_t* dist, fc_stats* stats) {
  using namespace fcd;
  fc::clear_error();
  if (!d_off || (E > 0 && !d_col) || (n > 0 && (!s || !t)) || n < 0)
    return fail(FC_ERR_INVALID, "fc_stcon_many_static: null argument or negative count");
  if (V < 1 || V > kStaticMaxV || E < 0 || E > kStaticMaxE)
    return fail(FC_ERR_INVALID, "fc_stcon_many_static: V=%d E=%lld outside the limits V <= %lld, E <= %lld", V,
                (long long)E, (long long)kStaticMaxV, (long long)kStaticMaxE);
  for (int32_t i = 0; i < n; ++i)
    if (s[i] < 0 || s[i] >= V || t[i] < 0 || t[i] >= V)
      return fail(FC_ERR_INVALID, "fc_stcon_many_static: query %d vertex out of range [0,%d)", i, V);
  if (stats) std::memset(stats, 0, sizeof(*stats));
  DeviceSlot* slot = nullptr;
  fc_status st = device_slot(&slot);
  if (st != FC_OK) return st;
  const VisitParams<int32_t> p = static_params(slot->ptr, d_off, d_col, V, E);
  const auto t_start = std::chrono::steady_clock::now();

  int32_t hs[kStaticQueries], ht[kStaticQueries], hi[kStaticQueries];
  VisitOut ho[64];
  int32_t i0 = 0;
  while (i0 < n) {
    // the next pass: up to kStaticQueries queries with s != t (s == t has
    // distance 0 and needs no visit)
    int32_t k = 0;
    for (; i0 < n && k < kStaticQueries; ++i0) {
      if (s[i0] == t[i0]) {
        if (connected) connected[i0] = 1;
        if (dist) dist[i0] = 0;
        continue;
      }
      hs[k] = s[i0];
      ht[k] = t[i0];
      hi[k] = i0;
      ++k;
    }
    if (k == 0) break;
    FC_HIP(hipMemcpy((void*)p.qs, hs, sizeof(int32_t) * (size_t)k, hipMemcpyHostToDevice));
    FC_HIP(hipMemcpy((void*)p.qt, ht, sizeof(int32_t) * (size_t)k, hipMemcpyHostToDevice));
    st = static_pass(*slot, p, k, hs, ht);
    if (st != FC_OK) return st;
    double wl = 0;
    for (int32_t j0 = 0; j0 < k; j0 += 64) {
      const int32_t m = std::min<int32_t>(64, k - j0);
      FC_HIP(hipMemcpy(ho, p.out + j0, sizeof(VisitOut) * (size_t)m, hipMemcpyDeviceToHost));
      for (int32_t j = 0; j < m; ++j) {
        const VisitOut& o = ho[j];
        if (o.status != 0) return fail(FC_ERR_CAPACITY, "visit kernel: frontier guard tripped (err=%d)", o.err);
        const int32_t i = hi[j0 + j];
        if (connected) connected[i] = o.dist >= 0 ? 1 : 0;
        if (dist) dist[i] = o.dist;
        wl += o.launches;
        if (stats) {
          stats->levels += o.levels;
          stats->depth += o.depth;
          stats->visited_vertices += o.visited_vertices;
          stats->visited_edges += o.visited_edges;
          stats->expanded_vertices += o.expanded_vertices;
          stats->expanded_edges += o.expanded_edges;
          stats->topdown_levels += o.topdown;
          stats->bottomup_levels += o.bottomup;
          stats->max_frontier = std::max(stats->max_frontier, o.max_frontier);
          stats->launches += o.launches;
        }
      }
    }
    slot->est_launches = 0.5 * slot->est_launches + 0.5 * (wl / k);
  }
  if (stats)
    stats->time_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
  return FC_OK;
}
