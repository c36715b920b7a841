// Visit engine for BFS and s–t connectivity on gfx950: one kernel launch per
// level, driven by a device-side phase machine in VisitCtl.
//
// Why launches and not a persistent kernel: on MI355X a grid-wide software
// barrier with the agent-scope release/acquire that cross-XCD visibility
// needs costs 16-35 us, a cooperative launch ~20 us, while a back-to-back
// ordinary launch costs ~2.4 us (1.6 us inside a hipGraph) and its boundary
// already orders every block's writes. So each launch does one level with
// all blocks. This replaces the reference's host-driven level loop with a
// 16-byte readback per level (BFS-4K/src/device/BFS_WorkEfficient.cu:81-118)
// and ST-CON's ~12 host round trips per level (st-conn/make_stcon.c:968-1187).
//
// No block votes and no block finalises. Every block loads the 1 KB control
// block at the start of a launch (one 16-byte load per lane of one wave) and
// derives the same step from it: whether the query is finished (met, empty
// frontier, error), the direction and the virtual-warp width. Block 0 writes
// the machine state of the next launch into the other half of a buffer
// indexed by launch parity, which no block of this launch reads, and writes
// the results of a finished query. A level's critical path is therefore the
// level's own memory traffic, with no atomic ticket and no serial chain of
// loads by a finalising block after it. The price is that a query's end is
// noticed at the start of the launch after its last level: an s–t query
// spends that launch expanding the next query's roots, so n queries of d
// levels cost n*d + 1 launches.
//
// s–t state words are 64-bit: (epoch << 32) | colour | level, and a word of
// another epoch is an unvisited vertex. Each query has its own epoch, so no
// query clears the state array: the next query starts in the same launch
// that sees the last one finish. Its roots are never claimed through the
// array; a probe that reaches the other colour's root proposes a meeting at
// the root's level 0.
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
//  * bottom-up (BFS): each wave owns 64 vertices; unvisited ones look for a
//    parent in the current frontier bitmap (direction-optimising BFS,
//    Beamer et al. SC'12).
// Frontier slots are reserved through LDS stages, one atomic per block, the
// in-house replacement of FrontierWrite::Write (BFS-4K/src/device/Util/GlobalWrite.cu:49-105)
// and of ST-CON's Thrust exclusive_scan / remove_copy compaction
// (st-conn/mythrustlib.cu:25-34,175-185), which cost O(V) per level.
//
// s–t claims take ONE atomicCAS per vertex (ST-CON k_stcon_binary_mask_unique_large,
// st-conn/cudakernel/stcon_kernel.cu:128-196). A CAS (or read) that shows the
// other colour is a meeting edge (v,u) and proposes the path length
// level(v)+1+level(u); the query ends after the level in which the first
// proposal appears, and the minimum proposal is the exact distance.
#include "visit.hpp"
#include "hubstage.hpp"

#include <hip/hip_runtime.h>

namespace fcd {

template <typename OffT>
__device__ __forceinline__ OffT bcast_off(OffT v, int src) {
  return sizeof(OffT) == 4 ? (OffT)bcast((int32_t)v, src) : (OffT)bcast64((int64_t)v, src);
}

__device__ __forceinline__ bool test_bit(const uint32_t* bm, int32_t v) {
  return (bm[v >> 5] >> (v & 31)) & 1u;
}

// Next-frontier appender of one wave. Normal-queue entries are staged in
// this wave's LDS buffer; a full buffer is published by the wave with one
// atomic, and at the end of the level the whole block publishes its buffers,
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
  unsigned long long* best_out;  // the query's best meeting candidate (device)
  bool bitmaps;                 // keep the visited and frontier bitmaps (BFS: bottom-up levels may follow)
  uint32_t epoch;               // s–t: tag of the current query
  int32_t root_own, root_other; // s–t: this colour's root and the other colour's
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

// Block-collective publication of both colours' staged entries (all threads,
// uniform control flow, at the end of the level). One barrier publishes every
// wave's counts; then four waves each take one list with one atomic, at the
// same time: wave 0 colour 0's normal list (with its degree sum and best
// meeting candidate), wave 1 colour 1's, waves 2 and 3 the two hub lists;
// after a second barrier every wave writes its entries.
struct BlockShared {
  uint32_t cnt[2][kWavesPerBlock];
  unsigned long long mf[2][kWavesPerBlock];
  unsigned long long best[2][kWavesPerBlock];
  uint32_t hcnt[2][kWavesPerBlock];
  unsigned long long hedges[2][kWavesPerBlock];
  uint32_t base[2];
  unsigned long long hbase[2];
};
static_assert(kWavesPerBlock >= 4, "one publishing wave per list");

template <typename OffT>
__device__ __forceinline__ void block_flush(Appender<OffT>& a, int ncol, BlockShared& sh, Ring* nxt0, Ring* nxt1) {
  const int w = (int)(threadIdx.x >> 6);
  const int lane = lane_id();
  const int own = a.c;  // this wave staged entries of its own colour only
  {
    const unsigned long long mf = (unsigned long long)wave_reduce_full_nz_u64<kWaveSum>(a.mf);
    const unsigned long long best = wave_reduce_full_nz_u64<kWaveMin>(a.best);
    const unsigned long long hd =
        wave_reduce_full_nz_u64<kWaveSum>((uint32_t)lane < a.h.cnt ? (unsigned long long)a.h.d[lane] : 0ull);
    if (lane == 0) {
      sh.cnt[own][w] = a.cnt;
      sh.mf[own][w] = mf;
      sh.best[own][w] = best;
      sh.hcnt[own][w] = a.h.cnt;
      sh.hedges[own][w] = hd;
      sh.cnt[own ^ 1][w] = 0;
      sh.mf[own ^ 1][w] = 0;
      sh.best[own ^ 1][w] = ~0ull;
      sh.hcnt[own ^ 1][w] = 0;
      sh.hedges[own ^ 1][w] = 0;
    }
  }
  __syncthreads();
  const int pc = w & 1;  // the list this wave publishes: colour pc, normal (w < 2) or hub (w >= 2)
  if (lane == 0 && w < 4 && pc < ncol) {
    Ring* nxt = pc ? nxt1 : nxt0;
    if (w < 2) {
      uint32_t tot = 0;
      unsigned long long mft = 0, bt = ~0ull;
      for (int i = 0; i < kWavesPerBlock; ++i) {
        const uint32_t k = sh.cnt[pc][i];
        sh.cnt[pc][i] = tot;
        tot += k;
        mft += sh.mf[pc][i];
        bt = sh.best[pc][i] < bt ? sh.best[pc][i] : bt;
      }
      sh.base[pc] = tot ? atomicAdd(&nxt->n, tot) : 0u;
      if (mft) atomicAdd((unsigned long long*)&nxt->mf, mft);
      if (bt != ~0ull) atomicMin(a.best_out, bt);
    } else {
      uint32_t k = 0;
      unsigned long long e = 0;
      for (int i = 0; i < kWavesPerBlock; ++i) {
        const uint32_t ki = sh.hcnt[pc][i];
        const unsigned long long ei = sh.hedges[pc][i];
        sh.hcnt[pc][i] = k;
        sh.hedges[pc][i] = e;
        k += ki;
        e += ei;
      }
      sh.hbase[pc] = k ? atomicAdd((unsigned long long*)&nxt->hub, ((unsigned long long)k << kHubEdgeBits) | e)
                       : 0ull;
    }
  }
  __syncthreads();
  const uint32_t base = sh.base[own] + sh.cnt[own][w];
  int32_t* q = a.p->q[own][a.nxt_par];
  for (uint32_t i = (uint32_t)lane; i < a.cnt; i += kWave) {
    const uint32_t slot = base + i;
    if (slot < a.p->qcap)
      q[slot] = a.buf[i];
    else
      atomicOr(&a.p->ctl->err, 1u);
  }
  if (a.h.cnt) {
    const unsigned long long hb = sh.hbase[own];
    hub_write(a.h, (uint32_t)(hb >> kHubEdgeBits) + sh.hcnt[own][w], (hb & kHubEdgeMask) + sh.hedges[own][w],
              a.p->hq[own][a.nxt_par], a.p->hb[own][a.nxt_par], a.p->hcap, &a.p->ctl->err);
  }
}

template <typename OffT>
__device__ __forceinline__ void propose(Appender<OffT>& a, int32_t dist, int32_t u) {
  const unsigned long long cand = ((unsigned long long)dist << 32) | (uint32_t)u;
  a.best = cand < a.best ? cand : a.best;
}

// Top-down probe of edge (v -> u) for colour c whose frontier is at level
// lvl, in three parts, so that a lane has the loads and the atomics of
// several edges in flight together:
//   probe_look   loads u's visited-bitmap word (BFS) or state word (s–t);
//   probe_issue  when that word leaves u unlabelled, issues the claiming
//                atomic together with the loads of u's offsets, so that a
//                claimed vertex's degree arrives with the atomic's result
//                instead of one round trip later;
//   probe_done   returns whether this lane claimed u (with its degree); for
//                s–t a meeting edge lowers the lane's best candidate.
template <typename OffT>
struct Probe {
  unsigned long long w;    // the looked-up word
  unsigned long long old;  // the atomic's result
  OffT ob, oe;             // u's offsets (when the atomic was issued)
  bool tried;              // the atomic was issued
};

template <typename OffT>
__device__ __forceinline__ Probe<OffT> probe_look(const Appender<OffT>& a, bool st, int32_t u, bool act) {
  const VisitParams<OffT>& p = *a.p;
  Probe<OffT> r;
  r.w = 0;
  r.old = 0;
  r.ob = r.oe = 0;
  r.tried = false;
  if (!act) return r;
  if (!st)
    r.w = p.visited[u >> 5];
  else if (u != a.root_own && u != a.root_other)
    r.w = p.sst[u];
  return r;
}

template <typename OffT>
__device__ __forceinline__ void probe_issue(const Appender<OffT>& a, bool st, int lvl, int32_t u, bool act,
                                            Probe<OffT>& r) {
  const VisitParams<OffT>& p = *a.p;
  if (!act) return;
  if (!st) {
    // BFS claims through the visited bitmap: one atomicOr per newly reached
    // vertex; the looked-up word filters most already-visited probes.
    const uint32_t bit = 1u << (u & 31);
    if ((uint32_t)r.w & bit) return;
    r.ob = p.off[u];
    r.oe = p.off[u + 1];
    r.old = atomicOr(p.visited + (u >> 5), bit);
    r.tried = true;
    return;
  }
  // A word of another epoch is unvisited.
  if (u == a.root_own || u == a.root_other || (uint32_t)(r.w >> 32) == a.epoch) return;
  r.ob = p.off[u];
  r.oe = p.off[u + 1];
  r.old = atomicCAS(&p.sst[u], r.w, st_word(a.epoch, a.c, lvl + 1));
  r.tried = true;
}

template <typename OffT>
__device__ __forceinline__ bool probe_done(Appender<OffT>& a, int32_t* parent, bool st, int lvl, int32_t v,
                                           int32_t u, bool act, const Probe<OffT>& r, int64_t& deg) {
  const VisitParams<OffT>& p = *a.p;
  deg = 0;
  if (!act) return false;
  if (!st) {
    if (!r.tried || ((uint32_t)r.old & (1u << (u & 31)))) return false;
    p.state[u] = lvl + 1;
    if (parent) parent[u] = v;
    deg = (int64_t)(r.oe - r.ob);
    return true;
  }
  if (u == a.root_own) return false;
  if (u == a.root_other) {
    propose(a, lvl + 1, u);
    return false;
  }
  unsigned long long cur = r.w;
  if (r.tried) {
    // The looked-up word may be stale, so a failed CAS retries with the value
    // it returns; within a launch the word only changes to this epoch, so
    // the second attempt settles it (three attempts in all).
    const unsigned long long mine = st_word(a.epoch, a.c, lvl + 1);
    cur = r.old;
    bool won = r.old == r.w;
    for (int k = 0; k < 2 && !won && (uint32_t)(cur >> 32) != a.epoch; ++k) {
      const unsigned long long old = atomicCAS(&p.sst[u], cur, mine);
      won = old == cur;
      cur = old;
    }
    if (won) {
      if (parent) parent[u] = v;
      deg = (int64_t)(r.oe - r.ob);
      return true;
    }
  }
  if ((uint32_t)(cur >> 32) != a.epoch) {
    atomicOr(&p.ctl->err, 4u);
    return false;
  }
  const int32_t w = (int32_t)(uint32_t)cur;
  if (((w & kColorBit) != 0) != (a.c != 0)) propose(a, lvl + 1 + (w & kLevelMask), u);
  return false;
}

template <typename OffT, int VW>
__device__ __forceinline__ void td_normal(Appender<OffT>& a, int32_t* parent, bool st, int lvl, int cur_par,
                                          uint32_t n, int gwave, int nwaves) {
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
      Probe<OffT> r0 = probe_look(a, st, u0, a0), r1 = probe_look(a, st, u1, a1);
      probe_issue(a, st, lvl, u0, a0, r0);
      probe_issue(a, st, lvl, u1, a1, r1);
      int64_t d0, d1;
      const bool c0 = probe_done(a, parent, st, lvl, v, u0, a0, r0, d0);
      const bool c1 = probe_done(a, parent, st, lvl, v, u1, a1, r1, d1);
      push_next(a, c0, u0, d0, a.bitmaps);
      push_next(a, c1, u1, d1, a.bitmaps);
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
__device__ __forceinline__ void td_hubs(Appender<OffT>& a, int32_t* parent, bool st, int lvl, int cur_par,
                                        uint64_t hubpk, int gwave, int nwaves) {
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
      Probe<OffT> r = probe_look(a, st, u, ok);
      probe_issue(a, st, lvl, u, ok, r);
      int64_t d;
      const bool cl = probe_done(a, parent, st, lvl, v, u, ok, r, d);
      push_next(a, cl, u, d, a.bitmaps);
    }
  }
}

// Level 0 of an s–t query: the root's adjacency [b, e) cut into 64-edge
// chunks over the waves of the grid (one edge per lane), whatever its degree.
template <typename OffT>
__device__ __forceinline__ void td_root(Appender<OffT>& a, int32_t* parent, int32_t r, OffT b, OffT e, int gwave,
                                        int nwaves) {
  const VisitParams<OffT>& p = *a.p;
  for (OffT k = b + (OffT)gwave * kWave; k < e; k += (OffT)nwaves * kWave) {
    const OffT j = k + (OffT)lane_id();
    const bool act = j < e;
    const int32_t u = act ? p.col[j] : 0;
    Probe<OffT> pr = probe_look(a, true, u, act);
    probe_issue(a, true, 0, u, act, pr);
    int64_t d;
    const bool cl = probe_done(a, parent, true, 0, r, u, act, pr, d);
    push_next(a, cl, u, d, false);
  }
}

// Candidates of degree <= kBuLane scan their own list; longer lists are
// scanned by the whole wave, 64 edges at a time, after the lane's own first
// round of eight.
constexpr int kBuLane = 128;

// Per-chunk loads of a bottom-up step (64 consecutive vertices of one wave):
// the unvisited flags and the offset pair. They are issued for the next
// chunk before the current chunk's columns are scanned, so a chunk costs
// about two dependent memory round trips instead of three.
template <typename OffT>
struct BuChunk {
  bool unvis;
  OffT b, e;
  uint32_t vword;  // the lane's visited word (lanes 0..31: 2w, 32..63: 2w + 1)
};

template <typename OffT>
__device__ __forceinline__ BuChunk<OffT> bu_load(const VisitParams<OffT>& p, uint32_t w) {
  const int lane = lane_id();
  const int32_t v = (int32_t)(w * 64u + (uint32_t)lane);
  const uint32_t wi = 2u * w + ((uint32_t)lane >> 5);
  BuChunk<OffT> c;
  c.vword = p.visited[wi];
  c.unvis = v < p.V && !((c.vword >> (lane & 31)) & 1u);
  c.b = c.e = 0;
  if (c.unvis) {
    c.b = p.off[v];
    c.e = p.off[v + 1];
  }
  return c;
}

// Bottom-up BFS step: every unvisited vertex looks for a neighbour in the
// current frontier bitmap.
template <typename OffT>
__device__ __forceinline__ void bu_step(Appender<OffT>& a, int32_t* parent, int lvl, const uint32_t* cfbm, int gwave,
                                        int nwaves) {
  const VisitParams<OffT>& p = *a.p;
  uint32_t* nfbm = a.nfbm;
  const int lane = lane_id();
  // Isolated vertices (the tail of a degree-ordered numbering) can have no
  // parent: their chunks are skipped (43% of the vertices of R-MAT s22).
  const uint32_t nw64 = ((uint32_t)p.v_active + 63u) >> 6;
  const int32_t mine = lvl + 1;
  // The wave's chunks are w = gwave + i * nwaves. Lane j first reads the
  // visited words of the j-th of the next 64 of them, so that chunks without
  // an unvisited vertex (most of them after the big bottom-up level) cost
  // nothing beyond that one round trip; the others are processed in order,
  // each loading the next one's visited bits and offsets while it works.
  for (uint64_t g0 = (uint64_t)gwave; g0 < nw64; g0 += (uint64_t)kWave * (uint64_t)nwaves) {
    const uint64_t cj = g0 + (uint64_t)lane * (uint64_t)nwaves;
    bool work = false;
    if (cj < nw64) {
      const uint64_t vis = ((uint64_t)p.visited[2 * cj + 1] << 32) | p.visited[2 * cj];
      const int64_t left = (int64_t)p.V - (int64_t)cj * 64;  // vertices of the chunk below V
      const uint64_t valid = left >= 64 ? ~0ull : (1ull << left) - 1ull;
      work = (~vis & valid) != 0;
    }
    uint64_t todo = ballot(work);
    if (!todo) continue;
    uint32_t w = (uint32_t)(g0 + (uint64_t)__builtin_ctzll(todo) * (uint64_t)nwaves);
    todo &= todo - 1;
    BuChunk<OffT> cur = bu_load(p, w);
    for (;;) {
      const bool more = todo != 0;
      uint32_t wn = 0;
      BuChunk<OffT> nx;
      if (more) {
        wn = (uint32_t)(g0 + (uint64_t)__builtin_ctzll(todo) * (uint64_t)nwaves);
        todo &= todo - 1;
        nx = bu_load(p, wn);
      }
      const int32_t v = (int32_t)(w * 64u + (uint32_t)lane);
      const bool cand = cur.unvis;
      const OffT b = cur.b, e = cur.e;
      const bool small = cand && (e - b) <= (OffT)kBuLane;
      int32_t par = -1;
      // A long list's first eight entries are scanned in the lane's own first
      // round too: they hold the parent of nearly every candidate (lists are in
      // id order, which puts the hubs first), so few long lists need the wave.
      const OffT lim = small ? e : (b + 8 < e ? b + 8 : e);
      if (cand) {
        // eight independent column loads per round, then the eight bitmap
        // words together, so a round costs two round trips whichever column
        // holds the parent; most candidates find one in the first round
        for (OffT j = b; j < lim && par < 0; j += 8) {
          int32_t u[8];
  #pragma unroll
          for (int k = 0; k < 8; ++k) u[k] = j + k < lim ? p.col[j + k] : -1;
          uint32_t fw[8];
  #pragma unroll
          for (int k = 0; k < 8; ++k) fw[k] = u[k] >= 0 ? cfbm[u[k] >> 5] : 0u;
  #pragma unroll
          for (int k = 7; k >= 0; --k)
            if (u[k] >= 0 && ((fw[k] >> (u[k] & 31)) & 1u)) par = u[k];
        }
      }
      uint64_t big = ballot(cand && !small && par < 0);
      while (big) {
        const int l = __ffsll((unsigned long long)big) - 1;
        big &= big - 1;
        const OffT bb = bcast_off<OffT>(lim, l), ee = bcast_off<OffT>(e, l);
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
            fp = bcast(u, __ffsll((unsigned long long)hm) - 1);
            break;
          }
        }
        if (lane == l) par = fp;
      }
      const bool claimed = cand && par >= 0;
      if (claimed) {
        p.state[v] = mine;
        if (parent) parent[v] = par;
      }
      // Bitmap words of these 64 vertices belong to this wave during a
      // bottom-up level, so they are updated without atomics (and the
      // prefetched words of the wave's next chunk are not changed by others).
      const uint64_t cm = ballot(claimed);
      // (the visited words came with the chunk's loads: no load waits here)
      if (lane == 0 && (uint32_t)cm) {
        p.visited[2u * w] = cur.vword | (uint32_t)cm;
        nfbm[2u * w] = (uint32_t)cm;
      }
      if (lane == 32 && (uint32_t)(cm >> 32)) {
        p.visited[2u * w + 1u] = cur.vword | (uint32_t)(cm >> 32);
        nfbm[2u * w + 1u] = (uint32_t)(cm >> 32);
      }
      push_next(a, claimed, v, claimed ? (int64_t)(e - b) : 0, false);
      if (!more) break;
      cur = nx;
      w = wn;
    }
  }
}

template <typename OffT>
__device__ __forceinline__ void td_dispatch(Appender<OffT>& a, int32_t* parent, bool st, int vw, int lvl, int cur_par,
                                            uint32_t n, int gwave, int nwaves) {
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
// Every block decides this at launch start: vw < ceil(mf / n) exactly when
// vw * n < mf, so no 64-bit division.
__device__ __forceinline__ int degree_width(uint32_t n, uint64_t mf_normal) {
  if (n == 0) return 4;
  int vw = 4;
  while (vw < 64 && (uint64_t)vw * n < mf_normal) vw <<= 1;
  return vw;
}

// Beamer's switch thresholds for direction-optimising BFS.
constexpr double kAlpha = 14.0;
constexpr double kBeta = 24.0;

__device__ __forceinline__ uint32_t ring_size(const Ring& r) { return r.n + (uint32_t)(r.hub >> kHubEdgeBits); }

__device__ __forceinline__ void zero_ring(Ring* r) { *r = Ring{0u, 0u, 0ull, 0ull}; }

// BFS reset (one launch): clear the bitmaps (and, on the caller's numbering,
// the levels and parents), and block 0 seeds the root.
template <typename OffT>
__device__ void bfs_reset(const VisitParams<OffT>& p, const VisitState& S, VisitState* nxt, int32_t* parent) {
  const int64_t gtid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t nthr = (int64_t)gridDim.x * kBlock;
  const uint32_t nwords = (((uint32_t)p.V + 63u) >> 6) * 2u;
  const int32_t src = S.src;
  if (!p.lazy_state) {
    for (int64_t v = gtid; v < p.V; v += nthr) {
      p.state[v] = v == src ? 0 : -1;
      if (parent) parent[v] = v == src ? (int32_t)v : -1;
    }
  }
  for (int64_t w = gtid; w < nwords; w += nthr) {
    const uint32_t bs = ((uint32_t)src >> 5) == (uint32_t)w ? 1u << (src & 31) : 0u;
    p.visited[w] = bs;
    p.fbm[0][0][w] = bs;
    p.fbm[0][1][w] = 0u;
    p.fbm[0][2][w] = 0u;
  }
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  VisitCtl* ctl = p.ctl;
  if (p.lazy_state) {
    // renumbered graph: only reached vertices' state and parent are written
    // (the visited bitmap marks them), so only the root's here
    p.state[src] = 0;
    if (parent) parent[src] = src;
  }
  const int64_t d = (int64_t)(p.off[src + 1] - p.off[src]);
  Ring r0{0u, 0u, 0ull, (uint64_t)d};
  if (d >= p.hub_threshold) {
    p.hq[0][0][0] = src;
    p.hb[0][0][0] = 0;
    r0.hub = (1ull << kHubEdgeBits) | (uint64_t)d;
  } else {
    p.q[0][0][0] = src;
    r0.n = 1u;
  }
  ctl->ring[0][0][0] = r0;
  zero_ring(&ctl->ring[0][0][1]);
  zero_ring(&ctl->ring[0][0][2]);
  VisitState n = S;
  n.phase = kPhaseLevel;
  n.lvl = 0;
  n.steps = n.td = n.bu = n.maxf = 0;
  n.launches = 1;
  n.bu_mode = 0;
  n.seen_edges = (uint64_t)d;
  n.visited_v = 1;
  n.visited_e = d;
  n.expv = n.expe = 0;
  *nxt = n;
}

template <typename OffT>
__global__ __launch_bounds__(kBlock) void phase_kernel(VisitCtl* ctl, int parity) {
  __shared__ VisitCtl h;
  __shared__ int32_t s_stage[kWavesPerBlock * kStageCap];
  __shared__ int32_t s_hub[2][kWavesPerBlock * kHubStage];
  __shared__ BlockShared s_sh;
  __shared__ OffT s_root[4];
  static_assert(sizeof(VisitCtl) == kWave * sizeof(uint4), "one 16-byte load per lane");
  if (threadIdx.x < kWave)
    reinterpret_cast<uint4*>(&h)[threadIdx.x] = reinterpret_cast<const uint4*>(ctl)[threadIdx.x];
  __syncthreads();
  const VisitParams<OffT>& p = *reinterpret_cast<const VisitParams<OffT>*>(h.params);
  const VisitState& S = h.st[parity];
  VisitState* nxt = &ctl->st[parity ^ 1];
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (S.phase == kPhaseIdle) {
    if (lead) nxt->phase = kPhaseIdle;
    return;
  }
  const bool st = h.mode != kModeBfs;
  int32_t* const parent = h.record_parents ? p.parent : nullptr;
  if (S.phase == kPhaseReset) {
    bfs_reset(p, S, nxt, parent);
    return;
  }

  // ---- level phase: every block takes the same decisions from S ----
  // (Scalars only below: a local copy of the state, or arrays indexed by the
  // colour, would be put in scratch memory.)
  int q = S.qi, L = S.lvl;
  int qp = q & 1;
  int32_t src = S.src, dst = S.dst;
  uint32_t epoch = S.epoch;
  bool start = st && L == 0;  // s–t: this launch expands the roots of query q
  const int ncol = st ? 2 : 1;
  Ring cur[2];
  cur[0] = h.ring[qp][0][L % 3];
  cur[1] = h.ring[qp][1][L % 3];
  const uint32_t nf0 = ring_size(cur[0]), nf1 = st ? ring_size(cur[1]) : 0u;
  const uint64_t mf01 = cur[0].mf + (st ? cur[1].mf : 0ull);
  // the level-L frontier (built by the previous launch) counted as labelled
  const bool counted = L > 0;
  const uint64_t seen = S.seen_edges + (counted ? mf01 : 0ull);
  if (!start) {
    // see whether the query is finished
    const bool empty = counted && (nf0 == 0 || (st && nf1 == 0));
    const unsigned long long best = h.best[qp];
    const bool runaway = S.steps > p.V + 2;
    if ((st && best != ~0ull) || empty || h.err || runaway) {
      if (lead) {
        const bool met = st && best != ~0ull;
        VisitOut o;
        o.status = (h.err || runaway) ? FC_CAPACITY_STATUS : 0;
        o.levels = S.steps;
        o.depth = met ? S.steps : 0;
        o.dist = met ? (int32_t)(best >> 32) : -1;
        o.meet = met ? (int32_t)(uint32_t)best : -1;
        o.topdown = S.td;
        o.bottomup = S.bu;
        o.max_frontier = S.maxf;
        o.visited_vertices = S.visited_v + (counted ? (int64_t)nf0 + (int64_t)nf1 : 0);
        o.visited_edges = S.visited_e + (counted ? (int64_t)mf01 : 0);
        o.expanded_vertices = S.expv;
        o.expanded_edges = S.expe;
        o.err = (int32_t)h.err;
        // a BFS spends one more launch noticing its empty frontier; an s–t
        // query notices its end in the launch that starts the next query
        o.launches = S.launches + (st ? 0 : 1);
        p.out[q] = o;
      }
      if (!st || q + 1 >= h.nq) {
        if (lead) nxt->phase = kPhaseIdle;
        return;
      }
      // the next s–t query starts in this launch
      ++q;
      qp = q & 1;
      src = S.nsrc;
      dst = S.ndst;
      epoch = S.epoch + 1;
      L = 0;
      start = true;
    }
  }

  const int64_t gtid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t nthr = (int64_t)gridDim.x * kBlock;
  const int gwave = (int)(blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6));
  const int nwaves = (int)(gridDim.x * kWavesPerBlock);
  int strat = kTopDown;
  int vw0 = 4, vw1 = 4;
  if (start) {
    if (threadIdx.x < 4) s_root[threadIdx.x] = p.off[(threadIdx.x < 2 ? src : dst) + (threadIdx.x & 1)];
    __syncthreads();
    const uint64_t ds = (uint64_t)(s_root[1] - s_root[0]), dt = (uint64_t)(s_root[3] - s_root[2]);
    if (lead) {
      // the new query's machine state after this launch (its level 0)
      nxt->phase = kPhaseLevel;
      nxt->qi = q;
      nxt->src = src;
      nxt->dst = dst;
      nxt->nsrc = q + 1 < h.nq ? p.qs[q + 1] : -1;
      nxt->ndst = q + 1 < h.nq ? p.qt[q + 1] : -1;
      nxt->lvl = 1;
      nxt->epoch = epoch;
      nxt->steps = 1;
      nxt->td = 2;
      nxt->bu = 0;
      nxt->maxf = 1;
      nxt->launches = 1;
      nxt->bu_mode = 0;
      nxt->seen_edges = ds + dt;
      nxt->visited_v = 2;
      nxt->visited_e = (int64_t)(ds + dt);
      nxt->expv = 2;
      nxt->expe = (int64_t)(ds + dt);
      p.sst[src] = st_word(epoch, 0, 0);
      p.sst[dst] = st_word(epoch, 1, 0);
      if (parent) {
        parent[src] = src;
        parent[dst] = dst;
      }
      // the next query's counters (this query's level-1 ring was cleared when
      // the previous query started), and this query's level-2 ring
      for (int c = 0; c < 2; ++c) {
        for (int k = 0; k < 3; ++k) zero_ring(&ctl->ring[qp ^ 1][c][k]);
        zero_ring(&ctl->ring[qp][c][2]);
      }
      ctl->best[qp ^ 1] = ~0ull;
    }
  } else {
    const double mu = (double)p.E - (double)seen;
    vw0 = h.vw_fixed > 0 ? h.vw_fixed : degree_width(cur[0].n, cur[0].mf - (cur[0].hub & kHubEdgeMask));
    if (st) vw1 = h.vw_fixed > 0 ? h.vw_fixed : degree_width(cur[1].n, cur[1].mf - (cur[1].hub & kHubEdgeMask));
    int bu = S.bu_mode;
    if (!st) {
      if (!bu) {
        if ((double)cur[0].mf > mu / kAlpha) bu = 1;
      } else if ((double)nf0 * kBeta < (double)p.V) {
        bu = 0;
      }
      strat = bu ? kBottomUp : kTopDown;
    }
    if (lead) {
      nxt->phase = kPhaseLevel;
      nxt->qi = q;
      nxt->src = src;
      nxt->dst = dst;
      nxt->nsrc = S.nsrc;
      nxt->ndst = S.ndst;
      nxt->lvl = L + 1;
      nxt->epoch = epoch;
      nxt->steps = S.steps + 1;
      nxt->td = S.td + (strat == kTopDown ? ncol : 0);
      nxt->bu = S.bu + (strat == kTopDown ? 0 : ncol);
      const int32_t mx = (int32_t)(nf0 > nf1 ? nf0 : nf1);
      nxt->maxf = mx > S.maxf ? mx : S.maxf;
      nxt->launches = S.launches + 1;
      nxt->bu_mode = bu;
      nxt->seen_edges = seen;
      nxt->visited_v = S.visited_v + (counted ? (int64_t)nf0 + (int64_t)nf1 : 0);
      nxt->visited_e = S.visited_e + (counted ? (int64_t)mf01 : 0);
      nxt->expv = S.expv + (int64_t)nf0 + (int64_t)nf1;
      nxt->expe = S.expe + (int64_t)mf01;
      zero_ring(&ctl->ring[qp][0][(L + 2) % 3]);
      if (st) zero_ring(&ctl->ring[qp][1][(L + 2) % 3]);
    }
    if (!st) {
      // the BFS frontier bitmap two levels ahead is free: clear it for the
      // next level's appends
      uint32_t* far = p.fbm[0][(L + 2) % 3];
      const uint32_t nwords = (((uint32_t)p.V + 63u) >> 6) * 2u;
      for (int64_t w = gtid; w < nwords; w += nthr) far[w] = 0u;
    }
  }

  // The waves are split between the colours in proportion to their frontier
  // edges, so that both colours' chains of dependent loads run at the same
  // time; each wave expands one colour and stages into that colour's lists.
  const int cur_par = L & 1;
  int w0 = nwaves;
  if (ncol == 2) {
    const double e0 = start ? (double)(s_root[1] - s_root[0]) : (double)cur[0].mf;
    const double e1 = start ? (double)(s_root[3] - s_root[2]) : (double)cur[1].mf;
    const int lo = e0 > 0 ? 1 : 0, hi = e1 > 0 ? nwaves - 1 : nwaves;
    const double f = e0 + e1 > 0 ? e0 / (e0 + e1) : 0.5;
    w0 = (int)((double)nwaves * f + 0.5);
    w0 = w0 < lo ? lo : (w0 > hi ? hi : w0);
  }
  const int cw = gwave < w0 ? 0 : 1;  // this wave's colour
  const int lw = cw ? gwave - w0 : gwave, nlw = cw ? nwaves - w0 : w0;
  Appender<OffT> a;
  a.p = &p;
  a.c = cw;
  a.nxt_par = cur_par ^ 1;
  a.nxt = &ctl->ring[qp][cw][(L + 1) % 3];
  a.nfbm = p.fbm[cw][(L + 1) % 3];
  a.best_out = &ctl->best[qp];
  a.bitmaps = !st;
  a.epoch = epoch;
  a.root_own = cw ? dst : src;
  a.root_other = cw ? src : dst;
  a.buf = s_stage + (threadIdx.x >> 6) * kStageCap;
  a.cnt = 0;
  a.mf = 0;
  a.best = ~0ull;
  a.h.u = &s_hub[0][(threadIdx.x >> 6) * kHubStage];
  a.h.d = &s_hub[1][(threadIdx.x >> 6) * kHubStage];
  a.h.cnt = 0;
  if (start) {
    td_root(a, parent, cw ? dst : src, s_root[2 * cw], s_root[2 * cw + 1], lw, nlw);
  } else if (strat == kTopDown) {
    // selects, not indexing: a local array indexed by a runtime value
    // would live in scratch memory
    const int vwc = cw ? vw1 : vw0;
    const uint32_t nc = cw ? cur[1].n : cur[0].n;
    const uint64_t hc = cw ? cur[1].hub : cur[0].hub;
    td_dispatch(a, parent, st, vwc, L, cur_par, nc, lw, nlw);
    td_hubs(a, parent, st, L, cur_par, hc, lw, nlw);
  } else {
    bu_step(a, parent, L, p.fbm[cw][L % 3], lw, nlw);
  }
  block_flush(a, ncol, s_sh, &ctl->ring[qp][0][(L + 1) % 3], &ctl->ring[qp][1][(L + 1) % 3]);
}

template <typename OffT>
hipError_t launch_phase(VisitCtl* ctl, int parity, int blocks, hipStream_t s) {
  hipLaunchKernelGGL(phase_kernel<OffT>, dim3(blocks), dim3(kBlock), 0, s, ctl, parity);
  return hipGetLastError();
}

template <typename OffT>
int phase_occupancy() {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, phase_kernel<OffT>, kBlock, 0) != hipSuccess) return 0;
  return nb;
}

template hipError_t launch_phase<int32_t>(VisitCtl*, int, int, hipStream_t);
template hipError_t launch_phase<int64_t>(VisitCtl*, int, int, hipStream_t);
template int phase_occupancy<int32_t>();
template int phase_occupancy<int64_t>();

}  // namespace fcd
