// Persistent level-loop visit kernel for BFS and s–t connectivity on gfx950.
//
// One cooperative launch runs a whole query: reset, every level, and the
// result write-back. Between levels all blocks meet at a bounded grid barrier
// (wave.hpp) instead of returning to the host, which replaces the reference's
// per-level launch + 16-byte readback loop (BFS-4K/src/device/BFS_WorkEfficient.cu:81-118)
// and ST-CON's ~12 host round trips per level (st-conn/make_stcon.c:968-1187).
//
// Frontier expansion (per colour, per step):
//  * top-down: normal-queue vertices are expanded by virtual warps of VW lanes
//    (4..64, chosen from the frontier's average degree), the reference's
//    BFS_KernelMainGLOB + EdgeVisit (BFS-4K/src/device/WorkEfficientKernel/BFS_WE_KernelMain.cu:2-53,
//    BFS_WE_Kernels1.cu:44-74) rebuilt for 64-lane waves; vertices of degree >=
//    hub_threshold are routed at claim time to a hub queue whose edges are
//    spread over the whole grid, the device-side work queue that replaces the
//    dynamic-parallelism child grids (BFS-4K/src/device/WorkEfficientKernel/BFS_WE_Dynamic.cu:4-50).
//  * bottom-up: each wave owns 64 vertices, unvisited ones look for a parent in
//    the current frontier bitmap (direction-optimising BFS, Beamer et al. SC'12).
// Frontier slots are reserved with one atomic per wave (64-bit ballot + mbcnt),
// the in-house replacement of FrontierWrite::Write (BFS-4K/src/device/Util/GlobalWrite.cu:49-105)
// and of ST-CON's Thrust exclusive_scan/remove_copy compaction
// (st-conn/mythrustlib.cu:25-34,175-185), which cost O(V) per level.
//
// s–t: the state word is colour | level; a vertex is claimed with ONE atomicCAS
// (ST-CON k_stcon_binary_mask_unique_large, st-conn/cudakernel/stcon_kernel.cu:128-196).
// A CAS that returns the other colour is a meeting edge (v,u) and proposes the
// path length level(v)+1+level(u); the visit ends after the step in which the
// first proposal appears and the minimum proposal is the exact distance.
#include "visit.hpp"

#include <hip/hip_runtime.h>

namespace fcd {

template <typename OffT>
__device__ __forceinline__ OffT shfl_off(OffT v, int src);
template <>
__device__ __forceinline__ int32_t shfl_off<int32_t>(int32_t v, int src) {
  return __shfl(v, src);
}
template <>
__device__ __forceinline__ int64_t shfl_off<int64_t>(int64_t v, int src) {
  return shfl64(v, src);
}

__device__ __forceinline__ bool test_bit(const uint32_t* bm, int32_t v) {
  return (bm[v >> 5] >> (v & 31)) & 1u;
}

// Probe edge (v -> u) for colour c whose frontier is at level lvl. Returns
// true when this lane claimed u for the next frontier.
template <typename OffT>
__device__ __forceinline__ bool probe_td(const VisitParams<OffT>& p, int c, int lvl, int32_t v,
                                         int32_t u, bool act) {
  if (!act) return false;
  if (p.mode == kModeBfs) {
    // BFS claims through the visited bitmap: one atomicOr per newly reached
    // vertex; the plain pre-load filters most already-visited probes.
    const uint32_t bit = 1u << (u & 31);
    uint32_t* w = p.visited + (u >> 5);
    if (*w & bit) return false;
    if (atomicOr(w, bit) & bit) return false;
    p.state[u] = lvl + 1;
    if (p.parent) p.parent[u] = v;
    return true;
  }
  const int32_t mine = (c ? kColorBit : 0) | (lvl + 1);
  int32_t cur = p.state[u];
  if (cur == -1) {
    cur = atomicCAS(&p.state[u], -1, mine);
    if (cur == -1) {
      atomicOr(p.visited + (u >> 5), 1u << (u & 31));
      if (p.parent) p.parent[u] = v;
      return true;
    }
  }
  if (((cur & kColorBit) != 0) != (c != 0)) {
    const unsigned long long cand = (unsigned long long)(lvl + 1 + (cur & kLevelMask));
    atomicMin(&p.ctl->best, (cand << 32) | (uint32_t)u);
  }
  return false;
}

// Wave-collective: put every claimed vertex into the next frontier of colour c
// (normal or hub queue by degree), optionally into the next frontier bitmap,
// and accumulate its degree into mf_acc.
template <typename OffT>
__device__ __forceinline__ void append_next(const VisitParams<OffT>& p, int c, int nxt_par, Ring* nxt,
                                            uint32_t* nfbm, bool claimed, int32_t u, uint64_t& mf_acc,
                                            bool set_bitmap) {
  int64_t d = 0;
  if (claimed) d = (int64_t)(p.off[u + 1] - p.off[u]);
  const bool hub = claimed && d >= p.hub_threshold;
  const bool norm = claimed && !hub;
  uint64_t m;
  const uint32_t slot = wave_reserve(&nxt->n, norm, &m);
  if (norm) {
    if (slot < p.qcap)
      p.q[c][nxt_par][slot] = u;
    else
      atomicOr(&p.ctl->err, 1u);
  }
  if (hub) {
    const unsigned long long old =
        atomicAdd((unsigned long long*)&nxt->hub, (1ull << kHubEdgeBits) | (unsigned long long)d);
    const uint32_t hs = (uint32_t)(old >> kHubEdgeBits);
    if (hs < p.hcap) {
      p.hq[c][nxt_par][hs] = u;
      p.hb[c][nxt_par][hs] = (int64_t)(old & kHubEdgeMask);
    } else {
      atomicOr(&p.ctl->err, 1u);
    }
  }
  if (claimed) {
    mf_acc += (uint64_t)d;
    if (set_bitmap) atomicOr(nfbm + (u >> 5), 1u << (u & 31));
  }
}

template <typename OffT, int VW>
__device__ void td_normal(const VisitParams<OffT>& p, int c, int lvl, int cur_par, int nxt_par, uint32_t n,
                          Ring* nxt, uint32_t* nfbm, uint64_t& mf_acc, int gwave, int nwaves) {
  constexpr int G = kWave / VW;
  const int lane = lane_id();
  const int grp = lane / VW, sub = lane % VW;
  const int32_t* q = p.q[c][cur_par];
  if (n > p.qcap) n = p.qcap;
  for (uint32_t base = (uint32_t)gwave * G; base < n; base += (uint32_t)nwaves * G) {
    const uint32_t i = base + (uint32_t)grp;
    int32_t v = 0;
    OffT b = 0, e = 0;
    if (i < n) {
      v = q[i];
      if ((uint32_t)v < (uint32_t)p.V) {
        b = p.off[v] + (OffT)sub;
        e = p.off[v + 1];
      } else {
        atomicOr(&p.ctl->err, 2u);
      }
    }
    while (ballot(b < e)) {
      const bool act = b < e;
      const int32_t u = act ? p.col[b] : 0;
      const bool cl = probe_td(p, c, lvl, v, u, act);
      append_next(p, c, nxt_par, nxt, nfbm, cl, u, mf_acc, true);
      b += VW;
    }
  }
}

// Edge-parallel expansion of the hub queue: the concatenated adjacency of all
// hub vertices is cut into chunks of kHubChunk edges handed out across waves.
constexpr int kHubChunk = 256;

template <typename OffT>
__device__ void td_hubs(const VisitParams<OffT>& p, int c, int lvl, int cur_par, int nxt_par, uint64_t hubpk,
                        Ring* nxt, uint32_t* nfbm, uint64_t& mf_acc, int gwave, int nwaves) {
  uint32_t nh = (uint32_t)(hubpk >> kHubEdgeBits);
  if (nh > p.hcap) nh = p.hcap;
  const uint64_t H = hubpk & kHubEdgeMask;
  if (nh == 0 || H == 0) return;
  const int lane = lane_id();
  const int32_t* hq = p.hq[c][cur_par];
  const int64_t* hb = p.hb[c][cur_par];
  for (uint64_t cb = (uint64_t)gwave * kHubChunk; cb < H; cb += (uint64_t)nwaves * kHubChunk) {
    uint32_t lo = 0, hi = nh;  // hb[0] == 0 <= cb; find the last slot with hb[slot] <= cb
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if ((uint64_t)hb[mid] <= cb)
        lo = mid;
      else
        hi = mid;
    }
    for (int k = 0; k < kHubChunk; k += kWave) {
      const uint64_t g = cb + (uint64_t)k + (uint64_t)lane;
      const bool act = g < H;
      int32_t v = 0, u = 0;
      if (act) {
        uint32_t s = lo;
        while (s + 1 < nh && (uint64_t)hb[s + 1] <= g) ++s;
        v = hq[s];
        bool good = (uint32_t)v < (uint32_t)p.V && (uint64_t)hb[s] <= g;
        OffT k = 0;
        if (good) {
          k = p.off[v] + (OffT)(g - (uint64_t)hb[s]);
          good = k < p.off[v + 1];
        }
        if (good)
          u = p.col[k];
        else {
          atomicOr(&p.ctl->err, 2u);
          v = -1;
        }
      }
      const bool ok = act && v >= 0 && (uint32_t)v < (uint32_t)p.V;
      const bool cl = probe_td(p, c, lvl, v, u, ok);
      append_next(p, c, nxt_par, nxt, nfbm, cl, u, mf_acc, true);
    }
  }
}

// Bottom-up step for colour c: every unvisited vertex looks for a neighbour in
// c's current frontier bitmap. In s–t modes the vertices of the other colour's
// current frontier are scanned too, so that a frontier-to-frontier edge is
// still seen as a meeting edge.
template <typename OffT>
__device__ void bu_step(const VisitParams<OffT>& p, int c, int lvl, const uint32_t* cfbm, const uint32_t* ofbm,
                        uint32_t* nfbm, int nxt_par, Ring* nxt, uint64_t& mf_acc, int gwave, int nwaves) {
  const int lane = lane_id();
  const uint32_t nw64 = ((uint32_t)p.V + 63u) >> 6;
  const int32_t mine = p.mode == kModeBfs ? lvl + 1 : ((c ? kColorBit : 0) | (lvl + 1));
  for (uint32_t w = (uint32_t)gwave; w < nw64; w += (uint32_t)nwaves) {
    const int32_t v = (int32_t)(w * 64u + (uint32_t)lane);
    const bool inr = v < p.V;
    const uint32_t wi = 2u * w + ((uint32_t)lane >> 5);
    const bool unvis = inr && !((p.visited[wi] >> (lane & 31)) & 1u);
    const bool oth = inr && ofbm && ((ofbm[wi] >> (lane & 31)) & 1u);
    const bool cand = unvis || oth;
    OffT b = 0, e = 0;
    if (cand) {
      b = p.off[v];
      e = p.off[v + 1];
    }
    const bool small = cand && (e - b) <= (OffT)32;
    int32_t par = -1;
    if (small) {
      for (OffT j = b; j < e; ++j) {
        const int32_t u = p.col[j];
        if (test_bit(cfbm, u)) {
          par = u;
          break;
        }
      }
    }
    uint64_t big = ballot(cand && !small);
    while (big) {
      const int l = __ffsll((unsigned long long)big) - 1;
      big &= big - 1;
      const OffT bb = shfl_off<OffT>(b, l), ee = shfl_off<OffT>(e, l);
      int32_t fp = -1;
      for (OffT j0 = bb; j0 < ee; j0 += kWave) {
        const OffT j = j0 + (OffT)lane;
        bool hit = false;
        int32_t u = 0;
        if (j < ee) {
          u = p.col[j];
          hit = test_bit(cfbm, u);
        }
        const uint64_t hm = ballot(hit);
        if (hm) {
          fp = __shfl(u, __ffsll((unsigned long long)hm) - 1);
          break;
        }
      }
      if (lane == l) par = fp;
    }
    const bool found = par >= 0;
    bool claimed = false;
    if (found && unvis) {
      claimed = true;
      p.state[v] = mine;
      if (p.parent) p.parent[v] = par;
    } else if (found && oth) {
      const int32_t sv = p.state[v];
      const unsigned long long cnd = (unsigned long long)(lvl + 1 + (sv & kLevelMask));
      atomicMin(&p.ctl->best, (cnd << 32) | (uint32_t)v);
    }
    // Bitmap words of these 64 vertices belong to this wave during a
    // bottom-up step, so they are updated without atomics.
    const uint64_t cm = ballot(claimed);
    if (lane == 0 && (uint32_t)cm) {
      p.visited[2u * w] |= (uint32_t)cm;
      nfbm[2u * w] = (uint32_t)cm;
    }
    if (lane == 32 && (uint32_t)(cm >> 32)) {
      p.visited[2u * w + 1u] |= (uint32_t)(cm >> 32);
      nfbm[2u * w + 1u] = (uint32_t)(cm >> 32);
    }
    append_next(p, c, nxt_par, nxt, nfbm, claimed, v, mf_acc, false);
  }
}

template <typename OffT>
__device__ void td_dispatch(const VisitParams<OffT>& p, int vw, int c, int lvl, int cur_par, int nxt_par, uint32_t n,
                            Ring* nxt, uint32_t* nfbm, uint64_t& mf_acc, int gwave, int nwaves) {
  switch (vw) {
    case 4: td_normal<OffT, 4>(p, c, lvl, cur_par, nxt_par, n, nxt, nfbm, mf_acc, gwave, nwaves); break;
    case 8: td_normal<OffT, 8>(p, c, lvl, cur_par, nxt_par, n, nxt, nfbm, mf_acc, gwave, nwaves); break;
    case 16: td_normal<OffT, 16>(p, c, lvl, cur_par, nxt_par, n, nxt, nfbm, mf_acc, gwave, nwaves); break;
    case 32: td_normal<OffT, 32>(p, c, lvl, cur_par, nxt_par, n, nxt, nfbm, mf_acc, gwave, nwaves); break;
    default: td_normal<OffT, 64>(p, c, lvl, cur_par, nxt_par, n, nxt, nfbm, mf_acc, gwave, nwaves); break;
  }
}

// Width for a top-down step: the smallest power of two >= the average degree
// of the normal-queue entries, clamped to [4, 64].
__device__ __forceinline__ int degree_width(uint32_t n, uint64_t mf_normal) {
  if (n == 0) return 4;
  const uint64_t avg = (mf_normal + n - 1) / n;
  int vw = 4;
  while (vw < 64 && (uint64_t)vw < avg) vw <<= 1;
  return vw;
}

// Scheduler MLP (12 -> 32 -> 32 -> 6, ReLU), evaluated by one thread; the
// output is a predicted log-cost per action. Parameter layout:
// W1[32][12] b1[32] W2[32][32] b2[32] W3[6][32] b3[6].
__device__ int mlp_argmin(const float* w, const float* x) {
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

__device__ __forceinline__ float flog2(double x) { return (float)__log2f((float)(x + 1.0)); }

// Frontier features of a scheduled s–t step (shared with the host tracer).
__device__ void step_features(float* f, const uint32_t n[2], const uint64_t mf[2], const uint64_t hubn[2], int lvl0,
                              int lvl1, int64_t V, int64_t E, uint64_t seen_edges) {
  const double mu = (double)E - (double)seen_edges;
  f[0] = flog2((double)n[0]) / 32.f;
  f[1] = flog2((double)mf[0]) / 40.f;
  f[2] = flog2((double)n[1]) / 32.f;
  f[3] = flog2((double)mf[1]) / 40.f;
  f[4] = flog2(mu > 0 ? mu : 0) / 40.f;
  f[5] = flog2((double)V) / 32.f;
  f[6] = flog2((double)E) / 40.f;
  f[7] = (float)(lvl0 < 64 ? lvl0 : 64) / 64.f;
  f[8] = (float)(lvl1 < 64 ? lvl1 : 64) / 64.f;
  f[9] = flog2(n[0] ? (double)mf[0] / n[0] : 0.0) / 20.f;
  f[10] = flog2(n[1] ? (double)mf[1] / n[1] : 0.0) / 20.f;
  f[11] = (float)(hubn[0] + hubn[1]) / (float)(n[0] + n[1] + 1);
}

// Beamer's switch thresholds for direction-optimising BFS.
constexpr double kAlpha = 14.0;
constexpr double kBeta = 24.0;

template <typename OffT>
__global__ __launch_bounds__(kBlock) void visit_kernel(VisitParams<OffT> p) {
  const uint32_t nb = gridDim.x;
  const int gwave = (int)(blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6));
  const int nwaves = (int)(nb * kWavesPerBlock);
  const int64_t gtid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t nthr = (int64_t)nb * kBlock;
  const uint32_t nwords = (((uint32_t)p.V + 63u) >> 6) * 2u;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  const bool st = p.mode != kModeBfs;
  VisitCtl* ctl = p.ctl;
  __shared__ int s_choice[3];

  int lvl[2] = {0, 0};
  if (!p.resume) {
    for (int64_t v = gtid; v < p.V; v += nthr) {
      int32_t s = -1;
      if (v == p.src)
        s = 0;
      else if (st && v == p.dst)
        s = kColorBit;
      p.state[v] = s;
      if (p.parent) p.parent[v] = s != -1 ? (int32_t)v : -1;
    }
    for (int64_t w = gtid; w < nwords; w += nthr) {
      const uint32_t bs = ((uint32_t)p.src >> 5) == (uint32_t)w ? 1u << (p.src & 31) : 0u;
      const uint32_t bt = (st && ((uint32_t)p.dst >> 5) == (uint32_t)w) ? 1u << (p.dst & 31) : 0u;
      p.visited[w] = bs | bt;
      p.fbm[0][0][w] = bs;
      p.fbm[0][1][w] = 0u;
      p.fbm[0][2][w] = 0u;
      if (st) {
        p.fbm[1][0][w] = bt;
        p.fbm[1][1][w] = 0u;
        p.fbm[1][2][w] = 0u;
      }
    }
    if (lead) {
      for (int c = 0; c < 2; ++c)
        for (int k = 0; k < 3; ++k) ctl->ring[c][k] = Ring{0u, 0u, 0ull, 0ull};
      for (int c = 0; c < (st ? 2 : 1); ++c) {
        const int32_t r = c ? p.dst : p.src;
        const int64_t d = (int64_t)(p.off[r + 1] - p.off[r]);
        Ring& R = ctl->ring[c][0];
        if (d >= p.hub_threshold) {
          p.hq[c][0][0] = r;
          p.hb[c][0][0] = 0;
          R.hub = (1ull << kHubEdgeBits) | (uint64_t)d;
        } else {
          p.q[c][0][0] = r;
          R.n = 1u;
        }
        R.mf = (uint64_t)d;
      }
      ctl->best = ~0ull;
      ctl->err = 0u;
    }
  } else {
    lvl[0] = p.lvl0;
    lvl[1] = p.lvl1;
  }
  int status = 0;
  if (!grid_sync(&ctl->bar, nb, p.spin_limit)) status = FC_TIMEOUT_STATUS;

  // Uniform bookkeeping (identical in every block).
  uint64_t seen_edges = 0;
  int64_t visited_v = 0;
  uint64_t visited_e = 0;
  if (p.resume) {
    seen_edges = ld_agent(&ctl->seen_edges);
    visited_v = ld_agent(&ctl->visited_v);
    visited_e = (uint64_t)ld_agent(&ctl->visited_e);
  } else {
    for (int c = 0; c < (st ? 2 : 1); ++c) {
      const Ring* R = &ctl->ring[c][0];
      seen_edges += ld_agent(&R->mf);
      visited_v += ld_agent(&R->n) + (ld_agent(&R->hub) >> kHubEdgeBits);
    }
    visited_e = seen_edges;
  }
  int steps = 0, td = 0, bu = 0, depth = 0, dist = -1, meet = -1;
  uint32_t maxf = 0;
  int64_t expv = 0, expe = 0;
  bool bu_mode = false;

  while (status == 0) {
    uint32_t cn[2] = {0u, 0u};
    uint64_t cmf[2] = {0ull, 0ull}, chub[2] = {0ull, 0ull};
    for (int c = 0; c < (st ? 2 : 1); ++c) {
      const Ring* R = &ctl->ring[c][lvl[c] % 3];
      cn[c] = ld_agent(&R->n);
      cmf[c] = ld_agent(&R->mf);
      chub[c] = ld_agent(&R->hub);
    }
    bool ex[2] = {false, false};
    int strat[2] = {kTopDown, kTopDown};
    int vw[2] = {0, 0};
    const double mu = (double)p.E - (double)seen_edges;
    if (p.mode == kModeBfs) {
      ex[0] = true;
      const uint32_t nf = cn[0] + (uint32_t)(chub[0] >> kHubEdgeBits);
      if (!bu_mode) {
        if ((double)cmf[0] > mu / kAlpha) bu_mode = true;
      } else if ((double)nf * kBeta < (double)p.V) {
        bu_mode = false;
      }
      strat[0] = bu_mode ? kBottomUp : kTopDown;
    } else if (p.mode == kModePlainSt) {
      ex[0] = ex[1] = true;
    } else {
      // Scheduled step: one side, chosen by the heuristic or the MLP.
      if (threadIdx.x == 0) {
        int side, sg, w;
        const uint32_t nf0 = cn[0] + (uint32_t)(chub[0] >> kHubEdgeBits);
        const uint32_t nf1 = cn[1] + (uint32_t)(chub[1] >> kHubEdgeBits);
        side = cmf[0] <= cmf[1] ? 0 : 1;
        sg = (double)cmf[side] > mu / kAlpha ? kBottomUp : kTopDown;
        w = degree_width(side ? cn[1] : cn[0], cmf[side] - (chub[side] & kHubEdgeMask)) >= 16 ? 64 : 8;
        if (p.policy == 2 && p.mlp) {
          float f[kFeatures];
          const uint32_t nn[2] = {nf0, nf1};
          const uint64_t hn[2] = {chub[0] >> kHubEdgeBits, chub[1] >> kHubEdgeBits};
          step_features(f, nn, cmf, hn, lvl[0], lvl[1], p.V, p.E, seen_edges);
          const int a = mlp_argmin(p.mlp, f);
          side = a / 3;
          sg = (a % 3) == 2 ? kBottomUp : kTopDown;
          w = (a % 3) == 1 ? 64 : 8;
        }
        if (p.force_side >= 0) side = p.force_side;
        if (p.force_strategy >= 0) {
          sg = p.force_strategy == 2 ? kBottomUp : kTopDown;
          w = p.force_strategy == 1 ? 64 : 8;
        }
        if (p.trace_feat && lead) {
          const uint32_t nn[2] = {nf0, nf1};
          const uint64_t hn[2] = {chub[0] >> kHubEdgeBits, chub[1] >> kHubEdgeBits};
          step_features(p.trace_feat, nn, cmf, hn, lvl[0], lvl[1], p.V, p.E, seen_edges);
        }
        s_choice[0] = side;
        s_choice[1] = sg;
        s_choice[2] = w;
      }
      __syncthreads();
      const int side = s_choice[0];
      ex[side] = true;
      strat[side] = s_choice[1];
      vw[side] = s_choice[2];
      __syncthreads();
    }

    // Clear the ring slot and frontier bitmap two steps ahead of each
    // expanding colour; nobody reads or writes them during this step.
    for (int c = 0; c < 2; ++c) {
      if (!ex[c]) continue;
      uint32_t* far = p.fbm[c][(lvl[c] + 2) % 3];
      for (int64_t w = gtid; w < nwords; w += nthr) far[w] = 0u;
      if (lead) ctl->ring[c][(lvl[c] + 2) % 3] = Ring{0u, 0u, 0ull, 0ull};
    }

    for (int c = 0; c < 2; ++c) {
      if (!ex[c]) continue;
      uint64_t mf_acc = 0;
      Ring* nxt = &ctl->ring[c][(lvl[c] + 1) % 3];
      const int cur_par = lvl[c] & 1, nxt_par = cur_par ^ 1;
      uint32_t* nfbm = p.fbm[c][(lvl[c] + 1) % 3];
      if (strat[c] == kTopDown) {
        int w = p.vw_fixed > 0 ? p.vw_fixed : vw[c];
        if (w == 0) w = degree_width(cn[c], cmf[c] - (chub[c] & kHubEdgeMask));
        td_dispatch(p, w, c, lvl[c], cur_par, nxt_par, cn[c], nxt, nfbm, mf_acc, gwave, nwaves);
        td_hubs(p, c, lvl[c], cur_par, nxt_par, chub[c], nxt, nfbm, mf_acc, gwave, nwaves);
      } else {
        const uint32_t* ofbm = st ? p.fbm[1 - c][lvl[1 - c] % 3] : nullptr;
        bu_step(p, c, lvl[c], p.fbm[c][lvl[c] % 3], ofbm, nfbm, nxt_par, nxt, mf_acc, gwave, nwaves);
      }
      const uint64_t s = wave_sum_u64(mf_acc);
      if (lane_id() == 0 && s) atomicAdd((unsigned long long*)&nxt->mf, (unsigned long long)s);
    }

    if (!grid_sync(&ctl->bar, nb, p.spin_limit)) {
      status = FC_TIMEOUT_STATUS;
      break;
    }
    ++steps;
    bool empty = false;
    for (int c = 0; c < 2; ++c) {
      if (!ex[c]) continue;
      const uint32_t nf = cn[c] + (uint32_t)(chub[c] >> kHubEdgeBits);
      expv += nf;
      expe += (int64_t)cmf[c];
      if (nf > maxf) maxf = nf;
      if (strat[c] == kTopDown)
        ++td;
      else
        ++bu;
      ++lvl[c];
      const Ring* R = &ctl->ring[c][lvl[c] % 3];
      const uint64_t nmf = ld_agent(&R->mf);
      const uint32_t nn = ld_agent(&R->n) + (uint32_t)(ld_agent(&R->hub) >> kHubEdgeBits);
      seen_edges += nmf;
      visited_e += nmf;
      visited_v += nn;
      if (nn == 0) empty = true;
    }
    if (st) {
      const unsigned long long best = ld_agent(&ctl->best);
      if (best != ~0ull) {
        dist = (int)(best >> 32);
        meet = (int)(uint32_t)best;
        depth = steps;
        break;
      }
    }
    if (empty) break;
    if (p.max_steps > 0 && steps >= p.max_steps) break;
    if (steps > p.V + 2) {  // cannot happen on a valid visit; guards against a runaway loop
      status = FC_CAPACITY_STATUS;
      break;
    }
  }

  if (lead) {
    const uint32_t err = ld_agent(&ctl->err);
    VisitOut o;
    o.status = status ? status : (err ? FC_CAPACITY_STATUS : 0);
    o.levels = steps;
    o.depth = depth;
    o.dist = dist;
    o.meet = meet;
    o.topdown = td;
    o.bottomup = bu;
    o.max_frontier = (int32_t)maxf;
    o.visited_vertices = visited_v;
    o.visited_edges = (int64_t)visited_e;
    o.expanded_vertices = expv;
    o.expanded_edges = expe;
    o.lvl_end[0] = lvl[0];
    o.lvl_end[1] = lvl[1];
    o.err = (int32_t)err;
    o.pad = 0;
    *p.out = o;
    ctl->seen_edges = seen_edges;
    ctl->visited_v = visited_v;
    ctl->visited_e = (int64_t)visited_e;
    __threadfence_system();
  }
}

template <typename OffT>
hipError_t launch_visit(const VisitParams<OffT>& p, int blocks, hipStream_t s) {
  void* args[] = {(void*)&p};
  return hipLaunchCooperativeKernel((const void*)visit_kernel<OffT>, dim3(blocks), dim3(kBlock), args, 0, s);
}

template <typename OffT>
int visit_occupancy() {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, visit_kernel<OffT>, kBlock, 0) != hipSuccess) return 0;
  return nb;
}

template hipError_t launch_visit<int32_t>(const VisitParams<int32_t>&, int, hipStream_t);
template hipError_t launch_visit<int64_t>(const VisitParams<int64_t>&, int, hipStream_t);
template int visit_occupancy<int32_t>();
template int visit_occupancy<int64_t>();

}  // namespace fcd
