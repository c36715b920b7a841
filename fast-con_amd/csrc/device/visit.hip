// Visit engine for BFS and s–t connectivity on gfx950: one kernel launch per
// phase (the reset of a query, or one level of it), driven by a device-side
// phase machine in VisitCtl.
//
// Why launches and not a persistent kernel: on MI355X a grid-wide software
// barrier with the agent-scope release/acquire that cross-XCD visibility
// needs costs 16-35 us, a cooperative launch ~20 us, while a back-to-back
// ordinary launch costs ~2.4 us (1.6 us inside a hipGraph) and its boundary
// already orders every block's writes. So each launch does one phase with all
// blocks, and the last block to finish (an atomic ticket, no waiting) takes
// the decisions for the next launch: direction, virtual-warp width and side,
// termination, result write-back, and the move to the next query. The host
// replays a captured hipGraph of identical launches; launches after the last
// query return at once. This replaces the reference's host-driven level loop
// with a 16-byte readback per level (BFS-4K/src/device/BFS_WorkEfficient.cu:81-118)
// and ST-CON's ~12 host round trips per level (st-conn/make_stcon.c:968-1187).
//
// Frontier expansion (per colour, per level):
//  * top-down: normal-queue vertices are expanded by virtual warps of VW lanes
//    (4..64, from the frontier's average degree): the reference's
//    BFS_KernelMainGLOB + EdgeVisit (BFS-4K/src/device/WorkEfficientKernel/BFS_WE_KernelMain.cu:2-53,
//    BFS_WE_Kernels1.cu:44-74) rebuilt for 64-lane waves. Vertices of degree >=
//    hub_threshold are routed at claim time to a hub queue whose concatenated
//    adjacency is cut into chunks spread over the whole grid: the device-side
//    work queue that replaces the dynamic-parallelism child grids
//    (BFS-4K/src/device/WorkEfficientKernel/BFS_WE_Dynamic.cu:4-50).
//  * bottom-up: each wave owns 64 vertices; unvisited ones look for a parent
//    in the current frontier bitmap (direction-optimising BFS, Beamer et al. SC'12).
// Frontier slots are reserved with one atomic per wave (64-bit ballot + mbcnt),
// the in-house replacement of FrontierWrite::Write (BFS-4K/src/device/Util/GlobalWrite.cu:49-105)
// and of ST-CON's Thrust exclusive_scan / remove_copy compaction
// (st-conn/mythrustlib.cu:25-34,175-185), which cost O(V) per level.
//
// s–t: the state word is colour | level; a vertex is claimed with ONE atomicCAS
// (ST-CON k_stcon_binary_mask_unique_large, st-conn/cudakernel/stcon_kernel.cu:128-196).
// A CAS (or read) that shows the other colour is a meeting edge (v,u) and
// proposes the path length level(v)+1+level(u); the visit ends after the level
// in which the first proposal appears, and the minimum proposal is the exact
// distance.
#include "visit.hpp"
#include "hubstage.hpp"

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

// Next-frontier appender of one wave. Normal-queue entries are staged in
// this wave's LDS buffer; a full buffer is published by the wave with one
// atomic, and at the end of the phase the whole block publishes its buffers,
// degree sums and best meeting candidate with one atomic each. A single
// frontier counter hit by every wave-iteration was the bottleneck of large
// levels (a contended word sustains only tens of millions of atomics/s).
constexpr uint32_t kStageCap = 1024;

template <typename OffT>
struct Appender {
  const VisitParams<OffT>* p;
  int c;
  int nxt_par;
  Ring* nxt;
  uint32_t* nfbm;
  bool bitmaps;                 // keep the visited and frontier bitmaps (bottom-up levels may follow)
  int32_t* buf;                 // this wave's LDS stage, kStageCap entries
  uint32_t cnt;                 // staged entries (wave-uniform)
  uint64_t mf;                  // per-lane degree sum of claimed vertices
  unsigned long long best;      // per-lane best meeting candidate (distance << 32 | vertex)
  HubStage h;                   // this wave's staged hub entries
};

template <typename OffT>
__device__ __forceinline__ void wave_flush(Appender<OffT>& a) {
  if (a.cnt == 0) return;
  __builtin_amdgcn_wave_barrier();
  uint32_t base = 0;
  if (lane_id() == 0) base = atomicAdd(&a.nxt->n, a.cnt);
  base = (uint32_t)__shfl((int)base, 0);
  int32_t* q = a.p->q[a.c][a.nxt_par];
  for (uint32_t i = (uint32_t)lane_id(); i < a.cnt; i += kWave) {
    const uint32_t slot = base + i;
    if (slot < a.p->qcap)
      q[slot] = a.buf[i];
    else
      atomicOr(&a.p->ctl->err, 1u);
  }
  __builtin_amdgcn_wave_barrier();
  a.cnt = 0;
}

// Wave-collective: add every claimed vertex (degree deg) to the next frontier
// of the appender's colour: normal entries and hub entries (degree >=
// hub_threshold) through their LDS stages. A per-lane atomic on the one hub
// counter serialised whole levels when thousands of hubs were claimed.
template <typename OffT>
__device__ __forceinline__ void push_next(Appender<OffT>& a, bool claimed, int32_t u, int64_t deg, bool set_bitmap) {
  const VisitParams<OffT>& p = *a.p;
  const bool hub = claimed && deg >= p.hub_threshold;
  const bool norm = claimed && !hub;
  const uint64_t m = ballot(norm);
  if (m) {
    const uint32_t k = (uint32_t)__popcll(m);
    if (a.cnt + k > kStageCap) wave_flush(a);
    if (norm) a.buf[a.cnt + lane_rank(m)] = u;
    a.cnt += k;
  }
  hub_push(a.h, hub, u, deg, (unsigned long long*)&a.nxt->hub, p.hq[a.c][a.nxt_par], p.hb[a.c][a.nxt_par], p.hcap,
           &p.ctl->err);
  if (claimed) {
    a.mf += (uint64_t)deg;
    if (set_bitmap) atomicOr(a.nfbm + (u >> 5), 1u << (u & 31));
  }
}

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, o);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), o);
    const unsigned long long w = ((unsigned long long)hi << 32) | lo;
    v = w < v ? w : v;
  }
  return v;
}

// Block-collective (all threads, uniform control flow): publish every wave's
// staged entries with one atomic, and the degree sum and best candidate.
struct BlockShared {
  uint32_t cnt[kWavesPerBlock];
  unsigned long long mf[kWavesPerBlock];
  unsigned long long best[kWavesPerBlock];
  uint32_t base;
  HubShared hub;
};

template <typename OffT>
__device__ __forceinline__ void block_flush(Appender<OffT>& a, BlockShared& sh) {
  const int w = (int)(threadIdx.x >> 6);
  const unsigned long long mf = (unsigned long long)wave_sum_u64(a.mf);
  const unsigned long long best = wave_min_u64(a.best);
  if (lane_id() == 0) {
    sh.cnt[w] = a.cnt;
    sh.mf[w] = mf;
    sh.best[w] = best;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
    unsigned long long mft = 0, bt = ~0ull;
    for (int i = 0; i < kWavesPerBlock; ++i) {
      const uint32_t c = sh.cnt[i];
      sh.cnt[i] = tot;
      tot += c;
      mft += sh.mf[i];
      bt = sh.best[i] < bt ? sh.best[i] : bt;
    }
    sh.base = tot ? atomicAdd(&a.nxt->n, tot) : 0u;
    if (mft) atomicAdd((unsigned long long*)&a.nxt->mf, mft);
    if (bt != ~0ull && bt < ld_agent(&a.p->ctl->best)) atomicMin(&a.p->ctl->best, bt);
  }
  __syncthreads();
  const uint32_t base = sh.base + sh.cnt[w];
  int32_t* q = a.p->q[a.c][a.nxt_par];
  for (uint32_t i = (uint32_t)lane_id(); i < a.cnt; i += kWave) {
    const uint32_t slot = base + i;
    if (slot < a.p->qcap)
      q[slot] = a.buf[i];
    else
      atomicOr(&a.p->ctl->err, 1u);
  }
  a.cnt = 0;
  a.mf = 0;
  a.best = ~0ull;
  __syncthreads();
  hub_block_flush(a.h, (unsigned long long*)&a.nxt->hub, a.p->hq[a.c][a.nxt_par], a.p->hb[a.c][a.nxt_par], a.p->hcap,
                  &a.p->ctl->err, sh.hub);
}

// Probe edge (v -> u) for colour c whose frontier is at level lvl. Returns
// true when this lane claimed u for the next frontier; a meeting edge lowers
// the lane's best candidate.
template <typename OffT>
__device__ __forceinline__ bool probe_td(Appender<OffT>& a, int32_t* parent, bool st, int lvl, int32_t v, int32_t u,
                                         bool act) {
  if (!act) return false;
  const VisitParams<OffT>& p = *a.p;
  if (!st) {
    // BFS claims through the visited bitmap: one atomicOr per newly reached
    // vertex; the plain pre-load filters most already-visited probes.
    const uint32_t bit = 1u << (u & 31);
    uint32_t* w = p.visited + (u >> 5);
    if (*w & bit) return false;
    if (atomicOr(w, bit) & bit) return false;
    p.state[u] = lvl + 1;
    if (parent) parent[u] = v;
    return true;
  }
  const int c = a.c;
  const int32_t mine = (c ? kColorBit : 0) | (lvl + 1);
  int32_t cur = p.state[u];
  if (cur == -1) {
    cur = atomicCAS(&p.state[u], -1, mine);
    if (cur == -1) {
      if (a.bitmaps) atomicOr(p.visited + (u >> 5), 1u << (u & 31));
      if (parent) parent[u] = v;
      return true;
    }
  }
  if (((cur & kColorBit) != 0) != (c != 0)) {
    const unsigned long long cand =
        ((unsigned long long)(lvl + 1 + (cur & kLevelMask)) << 32) | (uint32_t)u;
    a.best = cand < a.best ? cand : a.best;
  }
  return false;
}

template <typename OffT>
__device__ __forceinline__ int64_t degree_of(const VisitParams<OffT>& p, int32_t u) {
  return (int64_t)(p.off[u + 1] - p.off[u]);
}

template <typename OffT, int VW>
__device__ __forceinline__ void td_normal(Appender<OffT>& a, int32_t* parent, bool st, int lvl, int cur_par, uint32_t n, int gwave,
                          int nwaves) {
  const VisitParams<OffT>& p = *a.p;
  constexpr int G = kWave / VW;
  const int lane = lane_id();
  const int grp = lane / VW, sub = lane % VW;
  const int32_t* q = p.q[a.c][cur_par];
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
    // two columns per lane per round: both loads are issued before either
    // probe, halving the dependent round trips of long lists
    while (ballot(b < e)) {
      const bool a0 = b < e, a1 = b + VW < e;
      const int32_t u0 = a0 ? p.col[b] : 0;
      const int32_t u1 = a1 ? p.col[b + VW] : 0;
      const bool c0 = probe_td(a, parent, st, lvl, v, u0, a0);
      const bool c1 = probe_td(a, parent, st, lvl, v, u1, a1);
      push_next(a, c0, u0, c0 ? degree_of(p, u0) : 0, a.bitmaps);
      push_next(a, c1, u1, c1 ? degree_of(p, u1) : 0, a.bitmaps);
      b += 2 * VW;
    }
  }
}

// Edge-parallel expansion of the hub queue: the concatenated adjacency of all
// hub vertices is cut into chunks handed out across waves. A chunk is 256
// edges (four passes of one edge per lane) when there are enough edges for
// every wave to get a chunk of that size, else 128 or 64: when the hub edges
// are few (the second level of a visit from one root), a wave's chain of
// dependent loads per pass, not the edge count, sets the level's time, and
// smaller chunks spread the passes over more waves.
constexpr int kHubChunkMax = 256;

template <typename OffT>
__device__ __forceinline__ void td_hubs(Appender<OffT>& a, int32_t* parent, bool st, int lvl, int cur_par, uint64_t hubpk, int gwave,
                        int nwaves) {
  const VisitParams<OffT>& p = *a.p;
  uint32_t nh = (uint32_t)(hubpk >> kHubEdgeBits);
  if (nh > p.hcap) nh = p.hcap;
  const uint64_t H = hubpk & kHubEdgeMask;
  if (nh == 0 || H == 0) return;
  const int lane = lane_id();
  const int32_t* hq = p.hq[a.c][cur_par];
  const int64_t* hb = p.hb[a.c][cur_par];
  const uint64_t nw = (uint64_t)nwaves;
  const int chunk = H >= nw * kHubChunkMax ? kHubChunkMax : H >= nw * (kHubChunkMax / 2) ? kHubChunkMax / 2 : kWave;
  for (uint64_t cb = (uint64_t)gwave * (uint64_t)chunk; cb < H; cb += nw * (uint64_t)chunk) {
    // Last slot with hb[slot] <= cb (hb[0] == 0): a 64-ary search, each lane
    // testing one of 64 evenly spaced probes per round.
    uint32_t lo = 0, hi = nh;
    while (hi - lo > 1) {
      const uint32_t span = hi - lo;
      const uint32_t step = (span + 63u) / 64u;
      const uint32_t probe = lo + (uint32_t)lane * step;
      const bool le = probe < hi && (uint64_t)hb[probe] <= cb;
      const uint64_t m = ballot(le);  // a prefix of lanes (hb ascending); lane 0 always set
      const uint32_t k = 63u - (uint32_t)__builtin_clzll((unsigned long long)m);
      lo = lo + k * step;
      const uint32_t nhi = lo + step;
      hi = nhi < hi ? nhi : hi;
    }
    for (int k = 0; k < chunk; k += kWave) {
      const uint64_t g = cb + (uint64_t)k + (uint64_t)lane;
      const bool act = g < H;
      int32_t v = -1, u = 0;
      if (act) {
        uint32_t s = lo;
        while (s + 1 < nh && (uint64_t)hb[s + 1] <= g) ++s;
        v = hq[s];
        bool good = (uint32_t)v < (uint32_t)p.V && (uint64_t)hb[s] <= g;
        OffT k2 = 0;
        if (good) {
          k2 = p.off[v] + (OffT)(g - (uint64_t)hb[s]);
          good = k2 < p.off[v + 1];
        }
        if (good) {
          u = p.col[k2];
        } else {
          atomicOr(&p.ctl->err, 2u);
          v = -1;
        }
      }
      const bool ok = act && v >= 0;
      const bool cl = probe_td(a, parent, st, lvl, v, u, ok);
      push_next(a, cl, u, cl ? degree_of(p, u) : 0, a.bitmaps);
    }
  }
}

// Candidates of degree <= kBuLane scan their own list; longer lists are
// scanned by the whole wave, 64 edges at a time.
constexpr int kBuLane = 128;

// Bottom-up step for colour c: every unvisited vertex looks for a neighbour in
// c's current frontier bitmap. In s–t modes the vertices of the other colour's
// current frontier are scanned too, so that a frontier-to-frontier edge is
// still seen as a meeting edge.
// Per-chunk loads of a bottom-up step (64 consecutive vertices of one wave):
// the candidate flags and the offset pair. They are issued for the next
// chunk before the current chunk's columns are scanned, so a chunk costs
// about two dependent memory round trips instead of three.
template <typename OffT>
struct BuChunk {
  bool unvis, oth;
  OffT b, e;
};

template <typename OffT>
__device__ __forceinline__ BuChunk<OffT> bu_load(const VisitParams<OffT>& p, const uint32_t* ofbm, uint32_t w) {
  const int lane = lane_id();
  const int32_t v = (int32_t)(w * 64u + (uint32_t)lane);
  const bool inr = v < p.V;
  const uint32_t wi = 2u * w + ((uint32_t)lane >> 5);
  BuChunk<OffT> c;
  c.unvis = inr && !((p.visited[wi] >> (lane & 31)) & 1u);
  c.oth = inr && ofbm && ((ofbm[wi] >> (lane & 31)) & 1u);
  c.b = c.e = 0;
  if (c.unvis || c.oth) {
    c.b = p.off[v];
    c.e = p.off[v + 1];
  }
  return c;
}

template <typename OffT>
__device__ __forceinline__ void bu_step(Appender<OffT>& a, int32_t* parent, bool st, int lvl, const uint32_t* cfbm,
                        const uint32_t* ofbm, int gwave, int nwaves) {
  const VisitParams<OffT>& p = *a.p;
  const int c = a.c;
  uint32_t* nfbm = a.nfbm;
  const int lane = lane_id();
  // Isolated vertices (the tail of a degree-ordered numbering) can have no
  // parent: their chunks are skipped (43% of the vertices of R-MAT s22).
  const uint32_t nw64 = ((uint32_t)p.v_active + 63u) >> 6;
  const int32_t mine = st ? ((c ? kColorBit : 0) | (lvl + 1)) : lvl + 1;
  uint32_t w = (uint32_t)gwave;
  if (w >= nw64) return;
  BuChunk<OffT> cur = bu_load(p, ofbm, w);
  for (;;) {
    const uint32_t wn = w + (uint32_t)nwaves;
    const bool more = wn < nw64;
    BuChunk<OffT> nx;
    if (more) nx = bu_load(p, ofbm, wn);
    const int32_t v = (int32_t)(w * 64u + (uint32_t)lane);
    const bool unvis = cur.unvis, oth = cur.oth;
    const bool cand = unvis || oth;
    const OffT b = cur.b, e = cur.e;
    const bool small = cand && (e - b) <= (OffT)kBuLane;
    int32_t par = -1;
    if (small) {
      // eight independent column loads per round keep requests in flight;
      // most candidates find a parent in the first round
      for (OffT j = b; j < e && par < 0; j += 8) {
        const int32_t u0 = p.col[j];
        const int32_t u1 = j + 1 < e ? p.col[j + 1] : -1;
        const int32_t u2 = j + 2 < e ? p.col[j + 2] : -1;
        const int32_t u3 = j + 3 < e ? p.col[j + 3] : -1;
        const int32_t u4 = j + 4 < e ? p.col[j + 4] : -1;
        const int32_t u5 = j + 5 < e ? p.col[j + 5] : -1;
        const int32_t u6 = j + 6 < e ? p.col[j + 6] : -1;
        const int32_t u7 = j + 7 < e ? p.col[j + 7] : -1;
        if (test_bit(cfbm, u0)) par = u0;
        else if (u1 >= 0 && test_bit(cfbm, u1)) par = u1;
        else if (u2 >= 0 && test_bit(cfbm, u2)) par = u2;
        else if (u3 >= 0 && test_bit(cfbm, u3)) par = u3;
        else if (u4 >= 0 && test_bit(cfbm, u4)) par = u4;
        else if (u5 >= 0 && test_bit(cfbm, u5)) par = u5;
        else if (u6 >= 0 && test_bit(cfbm, u6)) par = u6;
        else if (u7 >= 0 && test_bit(cfbm, u7)) par = u7;
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
      if (parent) parent[v] = par;
    } else if (found && oth) {
      const int32_t sv = p.state[v];
      const unsigned long long cnd = ((unsigned long long)(lvl + 1 + (sv & kLevelMask)) << 32) | (uint32_t)v;
      a.best = cnd < a.best ? cnd : a.best;
    }
    // Bitmap words of these 64 vertices belong to this wave during a
    // bottom-up level, so they are updated without atomics (and the
    // prefetched words of the wave's next chunk are not changed by others).
    const uint64_t cm = ballot(claimed);
    if (lane == 0 && (uint32_t)cm) {
      p.visited[2u * w] |= (uint32_t)cm;
      nfbm[2u * w] = (uint32_t)cm;
    }
    if (lane == 32 && (uint32_t)(cm >> 32)) {
      p.visited[2u * w + 1u] |= (uint32_t)(cm >> 32);
      nfbm[2u * w + 1u] = (uint32_t)(cm >> 32);
    }
    push_next(a, claimed, v, claimed ? (int64_t)(e - b) : 0, false);
    if (!more) break;
    cur = nx;
    w = wn;
  }
}

template <typename OffT>
__device__ __forceinline__ void td_dispatch(Appender<OffT>& a, int32_t* parent, bool st, int vw, int lvl, int cur_par, uint32_t n,
                            int gwave, int nwaves) {
  switch (vw) {
    case 4: td_normal<OffT, 4>(a, parent, st, lvl, cur_par, n, gwave, nwaves); break;
    case 8: td_normal<OffT, 8>(a, parent, st, lvl, cur_par, n, gwave, nwaves); break;
    case 16: td_normal<OffT, 16>(a, parent, st, lvl, cur_par, n, gwave, nwaves); break;
    case 32: td_normal<OffT, 32>(a, parent, st, lvl, cur_par, n, gwave, nwaves); break;
    default: td_normal<OffT, 64>(a, parent, st, lvl, cur_par, n, gwave, nwaves); break;
  }
}

// Width for a top-down level: the smallest power of two >= the average degree
// of the normal-queue entries, clamped to [4, 64].
__device__ __forceinline__ int degree_width(uint32_t n, uint64_t mf_normal) {
  if (n == 0) return 4;
  const uint64_t avg = (mf_normal + n - 1) / n;
  int vw = 4;
  while (vw < 64 && (uint64_t)vw < avg) vw <<= 1;
  return vw;
}

// Beamer's switch thresholds for direction-optimising BFS.
constexpr double kAlpha = 14.0;
constexpr double kBeta = 24.0;

// Decision for the next level launch (one thread of the finalising block).
__device__ void decide(VisitCtl* ctl, int64_t V, int64_t E) {
  Decision d;
  d.ex[0] = d.ex[1] = 0;
  d.strat[0] = d.strat[1] = kTopDown;
  d.vw[0] = d.vw[1] = 8;
  uint32_t cn[2];
  uint64_t cmf[2], chub[2];
  for (int c = 0; c < 2; ++c) {
    const Ring* R = &ctl->ring[c][ctl->lvl[c] % 3];
    cn[c] = ld_agent(&R->n);
    cmf[c] = ld_agent(&R->mf);
    chub[c] = ld_agent(&R->hub);
  }
  const double mu = (double)E - (double)ctl->seen_edges;
  const int mode = ctl->mode;
  auto width = [&](int c) {
    return ctl->vw_fixed > 0 ? ctl->vw_fixed : degree_width(cn[c], cmf[c] - (chub[c] & kHubEdgeMask));
  };
  if (mode == kModeBfs) {
    d.ex[0] = 1;
    const uint64_t nf = (uint64_t)cn[0] + (chub[0] >> kHubEdgeBits);
    if (!ctl->bu_mode) {
      if ((double)cmf[0] > mu / kAlpha) ctl->bu_mode = 1;
    } else if ((double)nf * kBeta < (double)V) {
      ctl->bu_mode = 0;
    }
    d.strat[0] = ctl->bu_mode ? kBottomUp : kTopDown;
    d.vw[0] = width(0);
  } else {
    d.ex[0] = d.ex[1] = 1;
    d.vw[0] = width(0);
    d.vw[1] = width(1);
  }
  ctl->dec = d;
}

template <typename OffT>
__device__ void finalize_reset(const VisitParams<OffT>& p) {
  VisitCtl* ctl = p.ctl;
  const bool st = ctl->mode != kModeBfs;
  uint64_t seen = 0;
  for (int c = 0; c < 2; ++c)
    for (int k = 0; k < 3; ++k) ctl->ring[c][k] = Ring{0u, 0u, 0ull, 0ull};
  for (int c = 0; c < (st ? 2 : 1); ++c) {
    const int32_t r = c ? ctl->dst : ctl->src;
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
    seen += (uint64_t)d;
  }
  ctl->lvl[0] = ctl->lvl[1] = 0;
  ctl->steps = ctl->td = ctl->bu = ctl->maxf = 0;
  ctl->launches = 1;
  ctl->bu_mode = 0;
  ctl->best = ~0ull;
  ctl->err = 0u;
  ctl->seen_edges = seen;
  ctl->visited_v = st ? 2 : 1;
  ctl->visited_e = (int64_t)seen;
  ctl->expv = ctl->expe = 0;
  decide(ctl, p.V, p.E);
  ctl->phase = kPhaseLevel;
}

template <typename OffT>
__device__ void finalize_level(const VisitParams<OffT>& p, const Decision& d) {
  VisitCtl* ctl = p.ctl;
  const bool st = ctl->mode != kModeBfs;
  ctl->steps++;
  ctl->launches++;
  bool empty = false;
  for (int c = 0; c < 2; ++c) {
    if (!d.ex[c]) continue;
    const Ring* cur = &ctl->ring[c][ctl->lvl[c] % 3];
    const uint32_t nf = cur->n + (uint32_t)(cur->hub >> kHubEdgeBits);
    ctl->expv += nf;
    ctl->expe += (int64_t)cur->mf;
    if ((int32_t)nf > ctl->maxf) ctl->maxf = (int32_t)nf;
    if (d.strat[c] == kTopDown)
      ctl->td++;
    else
      ctl->bu++;
    ctl->lvl[c]++;
    const Ring* R = &ctl->ring[c][ctl->lvl[c] % 3];
    const uint64_t nmf = ld_agent(&R->mf);
    const uint32_t nn = ld_agent(&R->n) + (uint32_t)(ld_agent(&R->hub) >> kHubEdgeBits);
    ctl->seen_edges += nmf;
    ctl->visited_e += (int64_t)nmf;
    ctl->visited_v += nn;
    if (nn == 0) empty = true;
  }
  const unsigned long long best = ld_agent(&ctl->best);
  const uint32_t err = ld_agent(&ctl->err);
  const bool met = st && best != ~0ull;
  const bool runaway = ctl->steps > p.V + 2;
  if (met || empty || runaway || err) {
    VisitOut o;
    o.status = (err || runaway) ? FC_CAPACITY_STATUS : 0;
    o.levels = ctl->steps;
    o.depth = met ? ctl->steps : 0;
    o.dist = met ? (int32_t)(best >> 32) : -1;
    o.meet = met ? (int32_t)(uint32_t)best : -1;
    o.topdown = ctl->td;
    o.bottomup = ctl->bu;
    o.max_frontier = ctl->maxf;
    o.visited_vertices = ctl->visited_v;
    o.visited_edges = ctl->visited_e;
    o.expanded_vertices = ctl->expv;
    o.expanded_edges = ctl->expe;
    o.lvl_end[0] = ctl->lvl[0];
    o.lvl_end[1] = ctl->lvl[1];
    o.err = (int32_t)err;
    o.launches = ctl->launches;
    p.out[ctl->qi] = o;
    ctl->qi++;
    if (ctl->qi < ctl->nq) {
      ctl->src = p.qs[ctl->qi];
      ctl->dst = p.qt[ctl->qi];
      ctl->phase = kPhaseReset;
    } else {
      ctl->phase = kPhaseIdle;
    }
    return;
  }
  for (int c = 0; c < 2; ++c) ctl->ring[c][(ctl->lvl[c] + 1) % 3] = Ring{0u, 0u, 0ull, 0ull};
  decide(ctl, p.V, p.E);
}

template <typename OffT>
__global__ __launch_bounds__(kBlock) void phase_kernel(VisitParams<OffT> p) {
  VisitCtl* ctl = p.ctl;
  const int phase = ctl->phase;
  if (phase == kPhaseIdle) return;
  const uint32_t nb = gridDim.x;
  const int gwave = (int)(blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6));
  const int nwaves = (int)(nb * kWavesPerBlock);
  const int64_t gtid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t nthr = (int64_t)nb * kBlock;
  const uint32_t nwords = (((uint32_t)p.V + 63u) >> 6) * 2u;
  const bool st = ctl->mode != kModeBfs;
  // The plain s–t visit runs top-down only: it claims through state alone
  // and keeps neither the visited nor the frontier bitmaps.
  const bool bitmaps = ctl->mode != kModePlainSt;
  int32_t* const parent = ctl->record_parents ? p.parent : nullptr;
  __shared__ int s_last;
  __shared__ int32_t s_stage[kWavesPerBlock * kStageCap];
  __shared__ int32_t s_hub[2][kWavesPerBlock * kHubStage];
  __shared__ BlockShared s_sh;

  Decision d = ctl->dec;
  if (phase == kPhaseReset) {
    const int32_t src = ctl->src, dst = st ? ctl->dst : -1;
    if (st || !p.lazy_state) {
      for (int64_t v = gtid; v < p.V; v += nthr) {
        int32_t s = -1;
        if (v == src)
          s = 0;
        else if (v == dst)
          s = kColorBit;
        p.state[v] = s;
        if (parent) parent[v] = s != -1 ? (int32_t)v : -1;
      }
    } else if (gtid == 0) {
      // BFS on a renumbered graph: only reached vertices' state and parent
      // are written (the visited bitmap marks them), so only the root's here
      p.state[src] = 0;
      if (parent) parent[src] = src;
    }
    for (int64_t w = bitmaps ? gtid : (int64_t)nwords; w < nwords; w += nthr) {
      const uint32_t bs = ((uint32_t)src >> 5) == (uint32_t)w ? 1u << (src & 31) : 0u;
      const uint32_t bt = (dst >= 0 && ((uint32_t)dst >> 5) == (uint32_t)w) ? 1u << (dst & 31) : 0u;
      p.visited[w] = bs | bt;
      p.fbm[0][0][w] = bs;
      p.fbm[0][1][w] = 0u;
      p.fbm[0][2][w] = 0u;
      if (st) {  // the second colour's frontier bitmaps serve s–t modes only
        p.fbm[1][0][w] = bt;
        p.fbm[1][1][w] = 0u;
        p.fbm[1][2][w] = 0u;
      }
    }
  } else {
    const int lvl[2] = {ctl->lvl[0], ctl->lvl[1]};
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      if (!d.ex[c] || !bitmaps) continue;
      uint32_t* far = p.fbm[c][(lvl[c] + 2) % 3];
      for (int64_t w = gtid; w < nwords; w += nthr) far[w] = 0u;
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      if (!d.ex[c]) continue;
      const Ring cur = ctl->ring[c][lvl[c] % 3];
      const int cur_par = lvl[c] & 1;
      Appender<OffT> a;
      a.p = &p;
      a.c = c;
      a.nxt_par = cur_par ^ 1;
      a.nxt = &ctl->ring[c][(lvl[c] + 1) % 3];
      a.nfbm = p.fbm[c][(lvl[c] + 1) % 3];
      a.bitmaps = bitmaps;
      a.buf = s_stage + (threadIdx.x >> 6) * kStageCap;
      a.cnt = 0;
      a.mf = 0;
      a.best = ~0ull;
      a.h.u = &s_hub[0][(threadIdx.x >> 6) * kHubStage];
      a.h.d = &s_hub[1][(threadIdx.x >> 6) * kHubStage];
      a.h.cnt = 0;
      if (d.strat[c] == kTopDown) {
        td_dispatch(a, parent, st, d.vw[c], lvl[c], cur_par, cur.n, gwave, nwaves);
        td_hubs(a, parent, st, lvl[c], cur_par, cur.hub, gwave, nwaves);
      } else {
        const uint32_t* ofbm = st ? p.fbm[1 - c][lvl[1 - c] % 3] : nullptr;
        bu_step(a, parent, st, lvl[c], p.fbm[c][lvl[c] % 3], ofbm, gwave, nwaves);
      }
      block_flush(a, s_sh);
    }
  }

  // Arrival: every wave drains its own memory operations (including
  // non-returning atomics), then thread 0 takes a ticket; the block with the
  // last ticket finalises. The finaliser reads only counters that other
  // blocks updated with device-scope atomics (ring, best, err), all
  // performed before their tickets, so no cache writeback is needed here;
  // the plain stores of this launch (queues, states, bitmaps) are consumed
  // by the next launch, after the kernel boundary.
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    // Two-level ticket: the blocks of group b % kTicketGroups count on their
    // own line, and the last of each group counts on the launch ticket, so
    // no single address takes an atomic from every block.
    const uint32_t k = blockIdx.x % kTicketGroups;
    const uint32_t gsize = (nb - k + kTicketGroups - 1) / kTicketGroups;
    uint32_t* sub = &ctl->gticket[k * 32];
    const uint32_t t1 = __hip_atomic_fetch_add(sub, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    bool last = false;
    if (t1 == gsize - 1) {
      __hip_atomic_store(sub, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const uint32_t groups = nb < (uint32_t)kTicketGroups ? nb : (uint32_t)kTicketGroups;
      const uint32_t t = __hip_atomic_fetch_add(&ctl->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      last = t == groups - 1;
    }
    s_last = last;
  }
  __syncthreads();
  if (!s_last || threadIdx.x != 0) return;
  if (phase == kPhaseReset)
    finalize_reset(p);
  else
    finalize_level(p, d);
  __hip_atomic_store(&ctl->ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename OffT>
hipError_t launch_phase(const VisitParams<OffT>& p, int blocks, hipStream_t s) {
  hipLaunchKernelGGL(phase_kernel<OffT>, dim3(blocks), dim3(kBlock), 0, s, p);
  return hipGetLastError();
}

template <typename OffT>
int phase_occupancy() {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, phase_kernel<OffT>, kBlock, 0) != hipSuccess) return 0;
  return nb;
}

template hipError_t launch_phase<int32_t>(const VisitParams<int32_t>&, int, hipStream_t);
template hipError_t launch_phase<int64_t>(const VisitParams<int64_t>&, int, hipStream_t);
template int phase_occupancy<int32_t>();
template int phase_occupancy<int64_t>();

}  // namespace fcd
