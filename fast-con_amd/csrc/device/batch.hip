// Batched s–t connectivity: waves of 64 queries answered by one bit-parallel
// bidirectional visit (one 64-bit mask per vertex and side), so that the
// adjacency of hub vertices, which almost every R-MAT query touches, is read
// once per wave instead of once per query.
//
// Per vertex u and side x (0 = s side, 1 = t side):
//   rec[4u+2x]   queries whose side-x ball contains u
//   rec[4u+2x+1] masks ORed in during a top-down expansion, consumed by commit
//   F[x][p][u]   queries for which u is in side x's current frontier
// A step expands one side (heuristic / learned schedules) or both sides
// (plain schedule). With i and j the levels both sides had reached before the
// step, the balls of an unresolved query are disjoint, so its distance is
//   i+j+1  when a newly reached vertex of one side is already in the other
//          side's ball (an edge between the two frontiers), and
//   i+j+2  when a vertex is newly reached by both sides in the same step.
// A query is not connected when an expanded side gains no new vertex for it.
// These verdicts and distances equal the plain visit's (and the host BFS's).
//
// Execution follows the visit engine: one launch per phase, the last block
// to finish decides the next phase, the host replays captured launches.
// Only the masks a wave touched are cleared afterwards (dirty list), so a
// wave costs nothing proportional to V beyond the bottom-up steps.
#include "batch.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "engine.hpp"
#include "mlp.hpp"

namespace fcd {

using fc::fail;

constexpr uint32_t kBStage = 512;  // staged entries per wave and list
constexpr int kBLists = 3;

// ---------------------------------------------------------------------------
// LDS-staged appends (one wave stage per list; block publishes with one atomic)
// ---------------------------------------------------------------------------
struct Stage {
  int32_t* buf;
  uint32_t cnt;  // wave-uniform
};

__device__ __forceinline__ void stage_flush(Stage& s, uint32_t* counter, int32_t* list, uint32_t cap, uint32_t* err) {
  if (!s.cnt) return;
  __builtin_amdgcn_wave_barrier();
  uint32_t base = 0;
  if (lane_id() == 0) base = atomicAdd(counter, s.cnt);
  base = (uint32_t)__shfl((int)base, 0);
  for (uint32_t i = (uint32_t)lane_id(); i < s.cnt; i += kWave) {
    const uint32_t slot = base + i;
    if (slot < cap)
      list[slot] = s.buf[i];
    else
      atomicOr(err, 1u);
  }
  __builtin_amdgcn_wave_barrier();
  s.cnt = 0;
}

__device__ __forceinline__ void stage_push(Stage& s, bool want, int32_t v, uint32_t* counter, int32_t* list,
                                           uint32_t cap, uint32_t* err) {
  const uint64_t m = ballot(want);
  if (!m) return;
  const uint32_t k = (uint32_t)__popcll(m);
  if (s.cnt + k > kBStage) stage_flush(s, counter, list, cap, err);
  if (want) s.buf[s.cnt + lane_rank(m)] = v;
  s.cnt += k;
}

struct BShared {
  uint32_t cnt[kBLists][kWavesPerBlock];
  uint32_t base[kBLists];
  unsigned long long red[8][kWavesPerBlock];
  int last;
};

// Block-collective publication of the staged lists (all threads call).
__device__ __forceinline__ void stage_block_part(Stage& st, int l, uint32_t* counter, int32_t* list, uint32_t cap,
                                                 uint32_t* err, BShared& sh) {
  const int w = (int)(threadIdx.x >> 6);
  const uint32_t base = sh.base[l] + sh.cnt[l][w];
  for (uint32_t i = (uint32_t)lane_id(); i < st.cnt; i += kWave) {
    const uint32_t slot = base + i;
    if (slot < cap)
      list[slot] = st.buf[i];
    else
      atomicOr(err, 1u);
  }
  st.cnt = 0;
}

__device__ __forceinline__ void stages_block_flush(Stage& s0, Stage& s1, Stage& s2, int nl, uint32_t* const* counters,
                                                   int32_t* const* lists, const uint32_t* caps, uint32_t* err,
                                                   BShared& sh) {
  const int w = (int)(threadIdx.x >> 6);
  if (lane_id() == 0) {
    sh.cnt[0][w] = s0.cnt;
    sh.cnt[1][w] = s1.cnt;
    sh.cnt[2][w] = s2.cnt;
  }
  __syncthreads();
  if ((int)threadIdx.x < nl) {
    const int l = (int)threadIdx.x;
    uint32_t tot = 0;
    for (int i = 0; i < kWavesPerBlock; ++i) {
      const uint32_t c = sh.cnt[l][i];
      sh.cnt[l][i] = tot;
      tot += c;
    }
    sh.base[l] = tot ? atomicAdd(counters[l], tot) : 0u;
  }
  __syncthreads();
  stage_block_part(s0, 0, counters[0], lists[0], caps[0], err, sh);
  if (nl > 1) stage_block_part(s1, 1, counters[1], lists[1], caps[1], err, sh);
  if (nl > 2) stage_block_part(s2, 2, counters[2], lists[2], caps[2], err, sh);
  __syncthreads();
}

// Block reduction of up to 8 per-lane 64-bit values; op[k] = 0 sum, 1 or.
// Thread 0 then applies them to dst[k] with one atomic each (skipping zeros).
__device__ __forceinline__ void block_reduce_apply(unsigned long long* vals, const int* op, int nv, unsigned long long* const* dst,
                                   BShared& sh) {
  const int w = (int)(threadIdx.x >> 6);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if (k >= nv) break;
    unsigned long long v = vals[k];
    v = op[k] ? (unsigned long long)wave_or_u64(v) : (unsigned long long)wave_sum_u64(v);
    if (lane_id() == 0) sh.red[k][w] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 0; k < nv; ++k) {
      unsigned long long a = 0;
      for (int i = 0; i < kWavesPerBlock; ++i) a = op[k] ? (a | sh.red[k][i]) : (a + sh.red[k][i]);
      if (a) {
        if (op[k])
          atomicOr(dst[k], a);
        else
          atomicAdd(dst[k], a);
      }
    }
  }
  __syncthreads();
}

template <typename OffT>
__device__ __forceinline__ int64_t bdeg(const BatchParams<OffT>& p, int32_t u) {
  return (int64_t)(p.off[u + 1] - p.off[u]);
}

// Put a newly reached vertex u of side x into the next frontier: the normal
// list through the stage, or the hub list (with its edge base) by degree.
template <typename OffT>
__device__ __forceinline__ void push_frontier(const BatchParams<OffT>& p, Stage& st, int x, int nxt, bool want,
                                              int32_t u, int64_t deg) {
  const bool hub = want && deg >= p.hub_threshold;
  stage_push(st, want && !hub, u, &p.ctl->fl_n[x][nxt], p.fl[x][nxt], (uint32_t)p.V, &p.ctl->err);
  if (hub) {
    const unsigned long long old =
        atomicAdd(&p.ctl->hub[x][nxt], (1ull << kHubEdgeBits) | (unsigned long long)deg);
    const uint32_t hs = (uint32_t)(old >> kHubEdgeBits);
    if (hs < p.hcap) {
      p.hl[x][nxt][hs] = u;
      p.hb[x][nxt][hs] = (int64_t)(old & kHubEdgeMask);
    } else {
      atomicOr(&p.ctl->err, 1u);
    }
  }
}

// Top-down probe of edge (v -> u) for side x with frontier mask fm.
template <typename OffT>
__device__ __forceinline__ void probe(const BatchParams<OffT>& p, Stage& tst, int x, unsigned long long fm, int32_t u,
                                      bool act) {
  bool touched = false;
  if (act && fm) {
    unsigned long long* r = p.rec + 4 * (int64_t)u + 2 * x;
    const ulonglong2 vn = *reinterpret_cast<const ulonglong2*>(r);  // {visited, next}
    const unsigned long long m = fm & ~vn.x;
    // skip the atomic when every new bit is already in the next word
    if (m & ~vn.y) {
      const unsigned long long old = atomicOr(r + 1, m);
      if (old == 0) {
        const uint32_t bit = 1u << (u & 31);
        touched = !(atomicOr(&p.tb[u >> 5], bit) & bit);
      }
    }
  }
  stage_push(tst, touched, u, &p.ctl->touched_n, p.touched, (uint32_t)p.V, &p.ctl->err);
}

template <typename OffT, int VW>
__device__ __forceinline__ void expand_normal(const BatchParams<OffT>& p, Stage& tst, int x, int par, uint32_t n,
                              unsigned long long act, int gwave, int nwaves, unsigned long long& a_cnt,
                              unsigned long long& e_cnt) {
  constexpr int G = kWave / VW;
  const int lane = lane_id();
  const int grp = lane / VW, sub = lane % VW;
  const int32_t* list = p.fl[x][par];
  unsigned long long* F = p.F[x][par];
  if (n > (uint32_t)p.V) n = (uint32_t)p.V;
  for (uint32_t base = (uint32_t)gwave * G; base < n; base += (uint32_t)nwaves * G) {
    const uint32_t i = base + (uint32_t)grp;
    int32_t v = -1;
    OffT b = 0, e = 0;
    unsigned long long fm = 0;
    if (i < n) {
      v = list[i];
      if ((uint32_t)v < (uint32_t)p.V) {
        fm = F[v] & act;
        b = p.off[v] + (OffT)sub;
        e = p.off[v + 1];
        if (sub == 0 && fm) {
          a_cnt += 1;
          e_cnt += (unsigned long long)(e - b);
        }
      } else {
        atomicOr(&p.ctl->err, 2u);
        v = -1;
      }
    }
    // every lane of the group has read the mask before its leader clears it
    __builtin_amdgcn_wave_barrier();
    if (v >= 0 && sub == 0) F[v] = 0ull;
    if (!fm) e = b;
    while (ballot(b < e)) {
      const bool a0 = b < e, a1 = b + VW < e;
      const int32_t u0 = a0 ? p.col[b] : 0;
      const int32_t u1 = a1 ? p.col[b + VW] : 0;
      probe(p, tst, x, fm, u0, a0);
      probe(p, tst, x, fm, u1, a1);
      b += 2 * VW;
    }
  }
}

template <typename OffT>
__device__ __forceinline__ void expand_hubs(const BatchParams<OffT>& p, Stage& tst, int x, int par, unsigned long long hubpk,
                            unsigned long long act, int chunk, int gwave, int nwaves, unsigned long long& a_cnt,
                            unsigned long long& e_cnt) {
  uint32_t nh = (uint32_t)(hubpk >> kHubEdgeBits);
  if (nh > p.hcap) nh = p.hcap;
  const uint64_t H = hubpk & kHubEdgeMask;
  if (nh == 0 || H == 0) return;
  const int lane = lane_id();
  const int32_t* hl = p.hl[x][par];
  const int64_t* hb = p.hb[x][par];
  const unsigned long long* F = p.F[x][par];
  for (uint64_t cb = (uint64_t)gwave * (uint64_t)chunk; cb < H; cb += (uint64_t)nwaves * (uint64_t)chunk) {
    uint32_t lo = 0, hi = nh;
    while (hi - lo > 1) {
      const uint32_t span = hi - lo;
      const uint32_t step = (span + 63u) / 64u;
      const uint32_t pr = lo + (uint32_t)lane * step;
      const bool le = pr < hi && (uint64_t)hb[pr] <= cb;
      const uint64_t m = ballot(le);
      const uint32_t k = 63u - (uint32_t)__builtin_clzll((unsigned long long)m);
      lo = lo + k * step;
      const uint32_t nhi = lo + step;
      hi = nhi < hi ? nhi : hi;
    }
    for (int k = 0; k < chunk; k += kWave) {
      const uint64_t g = cb + (uint64_t)k + (uint64_t)lane;
      bool act_e = g < H;
      int32_t u = 0;
      unsigned long long fm = 0;
      if (act_e) {
        uint32_t s = lo;
        while (s + 1 < nh && (uint64_t)hb[s + 1] <= g) ++s;
        const int32_t v = hl[s];
        bool good = (uint32_t)v < (uint32_t)p.V && (uint64_t)hb[s] <= g;
        OffT k2 = 0;
        if (good) {
          k2 = p.off[v] + (OffT)(g - (uint64_t)hb[s]);
          good = k2 < p.off[v + 1];
        }
        if (good) {
          u = p.col[k2];
          fm = F[v] & act;
          if (fm) {
            e_cnt += 1;
            if ((uint64_t)hb[s] == g) a_cnt += 1;  // first edge of a hub with live queries
          }
        } else {
          atomicOr(&p.ctl->err, 2u);
          act_e = false;
        }
      }
      probe(p, tst, x, fm, u, act_e);
    }
  }
}

template <typename OffT>
__device__ __forceinline__ void expand_side(const BatchParams<OffT>& p, Stage& tst, int x, int vw, int chunk, unsigned long long act,
                            int gwave, int nwaves, unsigned long long& a_cnt, unsigned long long& e_cnt) {
  const BatchCtl* ctl = p.ctl;
  const int par = ctl->cur[x];
  const uint32_t n = ctl->fl_n[x][par];
  switch (vw) {
    case 4: expand_normal<OffT, 4>(p, tst, x, par, n, act, gwave, nwaves, a_cnt, e_cnt); break;
    case 8: expand_normal<OffT, 8>(p, tst, x, par, n, act, gwave, nwaves, a_cnt, e_cnt); break;
    case 16: expand_normal<OffT, 16>(p, tst, x, par, n, act, gwave, nwaves, a_cnt, e_cnt); break;
    case 32: expand_normal<OffT, 32>(p, tst, x, par, n, act, gwave, nwaves, a_cnt, e_cnt); break;
    default: expand_normal<OffT, 64>(p, tst, x, par, n, act, gwave, nwaves, a_cnt, e_cnt); break;
  }
  expand_hubs(p, tst, x, par, ctl->hub[x][par], act, chunk, gwave, nwaves, a_cnt, e_cnt);
}

// Commit one touched vertex (both sides); appends go through stages 0/1
// (frontier lists of s and t) and 2 (dirty list).
template <typename OffT>
__device__ __forceinline__ void commit_vertex(const BatchParams<OffT>& p, Stage& s0, Stage& s1, Stage& s2, bool valid,
                                              int32_t u,
                                              unsigned long long act, const int* nxt, unsigned long long* red) {
  unsigned long long nS = 0, nT = 0, vS = 0, vT = 0;
  int64_t deg = 0;
  if (valid) {
    const uint32_t bit = 1u << (u & 31);
    atomicAnd(&p.tb[u >> 5], ~bit);
    ulonglong2* r = reinterpret_cast<ulonglong2*>(p.rec + 4 * (int64_t)u);
    const ulonglong2 a = r[0], b = r[1];  // {visS, NS}, {visT, NT}
    vS = a.x;
    vT = b.x;
    nS = a.y & ~vS & act;
    nT = b.y & ~vT & act;
    red[0] |= (nS & vT) | (nT & vS);  // odd meets
    red[1] |= nS & nT;                 // even meets
    red[2] |= nS;
    red[3] |= nT;
    if (a.y) r[0] = make_ulonglong2(vS | nS, 0ull);
    if (b.y) r[1] = make_ulonglong2(vT | nT, 0ull);
    if (nS) p.F[0][nxt[0]][u] = nS;
    if (nT) p.F[1][nxt[1]][u] = nT;
    if (nS | nT) deg = bdeg(p, u);
    if (nS) red[4] += (unsigned long long)deg;
    if (nT) red[5] += (unsigned long long)deg;
  }
  push_frontier(p, s0, 0, nxt[0], nS != 0, u, deg);
  push_frontier(p, s1, 1, nxt[1], nT != 0, u, deg);
  stage_push(s2, ((vS | vT) == 0) && ((nS | nT) != 0), u, &p.ctl->dirty_n, p.dirty, (uint32_t)p.V + 256u,
             &p.ctl->err);
}

// Bottom-up step for side x: every vertex still needed by some query ORs the
// current frontier masks of its neighbours, stopping once it has them all.
template <typename OffT>
__device__ __forceinline__ void bottom_up(const BatchParams<OffT>& p, Stage& s0, Stage& s1, Stage& s2, int x,
                                          unsigned long long act, int gwave, int nwaves,
                          unsigned long long* red, unsigned long long& a_cnt, unsigned long long& e_cnt) {
  const BatchCtl* ctl = p.ctl;
  const int cur = ctl->cur[x], nxt = cur ^ 1;
  const unsigned long long* F = p.F[x][cur];
  const int lane = lane_id();
  const int64_t nthr = (int64_t)nwaves * kWave;
  for (int64_t base = (int64_t)gwave * kWave; base < p.V; base += nthr) {
    const int32_t u = (int32_t)(base + lane);
    unsigned long long acc = 0, need = 0, vx = 0, vo = 0;
    OffT b = 0, e = 0;
    if (u < p.V) {
      vx = p.rec[4 * (int64_t)u + 2 * x];
      need = act & ~vx;
      if (need) {
        b = p.off[u];
        e = p.off[u + 1];
      }
    }
    if (b < e) {
      a_cnt += 1;
      for (OffT j = b; j < e; j += 4) {
        const int32_t c0 = p.col[j];
        const int32_t c1 = j + 1 < e ? p.col[j + 1] : -1;
        const int32_t c2 = j + 2 < e ? p.col[j + 2] : -1;
        const int32_t c3 = j + 3 < e ? p.col[j + 3] : -1;
        acc |= F[c0];
        if (c1 >= 0) acc |= F[c1];
        if (c2 >= 0) acc |= F[c2];
        if (c3 >= 0) acc |= F[c3];
        e_cnt += (unsigned long long)((j + 4 <= e) ? 4 : (e - j));
        if ((acc & need) == need) break;
      }
    }
    const unsigned long long n = acc & need;
    int64_t deg = 0;
    if (n) {
      vo = p.rec[4 * (int64_t)u + 2 * (1 - x)];
      p.rec[4 * (int64_t)u + 2 * x] = vx | n;
      p.F[x][nxt][u] = n;
      red[0] |= n & vo;
      red[2 + x] |= n;
      deg = (int64_t)(e - b);
      red[4 + x] += (unsigned long long)deg;
    }
    if (x == 0)
      push_frontier(p, s0, 0, nxt, n != 0, u, deg);
    else
      push_frontier(p, s1, 1, nxt, n != 0, u, deg);
    stage_push(s2, n != 0 && vx == 0 && vo == 0, u, &p.ctl->dirty_n, p.dirty, (uint32_t)p.V + 256u, &p.ctl->err);
  }
}

// Features of the pending step for the learned scheduler.
__device__ __forceinline__ void batch_features(BatchCtl* ctl, int64_t V, int64_t E) {
  uint32_t n[2];
  unsigned long long mf[2], hn[2];
  for (int x = 0; x < 2; ++x) {
    const int c = ctl->cur[x];
    hn[x] = ctl->hub[x][c] >> kHubEdgeBits;
    n[x] = ctl->fl_n[x][c] + (uint32_t)hn[x];
    mf[x] = ctl->fl_mf[x][c];
  }
  const uint64_t mfs[2] = {mf[0], mf[1]};
  const uint64_t hns[2] = {hn[0], hn[1]};
  step_features(ctl->feat, n, mfs, hns, ctl->lvl[0], ctl->lvl[1], V, E, 0);
}

__device__ __forceinline__ int bdegree_width(uint32_t n, uint64_t mf) {
  if (n == 0) return 4;
  const uint64_t avg = (mf + n - 1) / n;
  int vw = 4;
  while (vw < 64 && (uint64_t)vw < avg) vw <<= 1;
  return vw;
}

__device__ __forceinline__ void batch_apply_action(BatchDecision& d, int a) {
  d.sides = (a / 3) == 0 ? kSideS : kSideT;
  d.strat = (a % 3) == 2 ? kBottomUp : kTopDown;
  d.vw = (a % 3) == 1 ? 64 : 8;
  d.action = a;
}

// Next step of the current wave (one thread of the finalising block).
__device__ __forceinline__ void batch_decide(BatchCtl* ctl, const float* mlp, int64_t V, int64_t E) {
  BatchDecision d;
  unsigned long long mf[2];
  uint32_t nn[2];
  for (int x = 0; x < 2; ++x) {
    const int c = ctl->cur[x];
    mf[x] = ctl->fl_mf[x][c];
    nn[x] = ctl->fl_n[x][c];
  }
  if (ctl->policy == 0) {
    d.sides = kSideBoth;
    d.strat = kTopDown;
    d.action = -1;
    const uint32_t n = nn[0] + nn[1];
    const unsigned long long m = (mf[0] - (ctl->hub[0][ctl->cur[0]] & kHubEdgeMask)) +
                                 (mf[1] - (ctl->hub[1][ctl->cur[1]] & kHubEdgeMask));
    d.vw = bdegree_width(n, m);
  } else {
    const int side = mf[0] <= mf[1] ? 0 : 1;
    // A bit-parallel bottom-up vertex can stop only once it has found every
    // live query's bit, so it scans nearly its whole list: worth it only when
    // the top-down frontier would touch most of the graph anyway.
    const bool bu = (double)mf[side] * 2.0 > (double)E;
    const int w = bdegree_width(nn[side], mf[side] - (ctl->hub[side][ctl->cur[side]] & kHubEdgeMask));
    int a = side * 3 + (bu ? 2 : (w >= 16 ? 1 : 0));
    batch_features(ctl, V, E);
    if (ctl->policy == 2 && mlp) a = mlp_argmin(mlp, ctl->feat);
    batch_apply_action(d, a);
    if (d.strat == kTopDown) d.vw = w;
    if (ctl->policy == 2 && (a % 3) != 2) d.vw = (a % 3) == 1 ? 64 : 8;
  }
  if (ctl->vw_fixed > 0 && d.strat == kTopDown) d.vw = ctl->vw_fixed;
  ctl->dec = d;
}

template <typename OffT>
__device__ __forceinline__ void batch_finalize(const BatchParams<OffT>& p, int phase, const BatchDecision& d) {
  BatchCtl* ctl = p.ctl;
  const int wave0 = ctl->wave * 64;
  switch (phase) {
    case kBInit: {
      unsigned long long act = 0;
      for (int q = 0; q < 64; ++q) {
        const int qi = wave0 + q;
        if (qi >= ctl->nq) break;
        if (p.qs[qi] == p.qt[qi])
          p.dist[qi] = 0;
        else
          act |= 1ull << q;
      }
      ctl->act = act;
      ctl->lvl[0] = ctl->lvl[1] = 0;
      for (int x = 0; x < 2; ++x) {
        const int c = ctl->cur[x];
        ctl->fl_mf[x][c] = ld_agent(&ctl->fl_mf[x][c]);
      }
      if (!act) {
        ctl->phase = kBCleanup;
        return;
      }
      batch_decide(ctl, p.mlp, p.V, p.E);
      ctl->phase = ctl->dec.strat == kBottomUp ? kBBottomUp : kBExpand;
      return;
    }
    case kBExpand:
      ctl->phase = kBCommit;
      return;
    case kBCommit:
    case kBBottomUp: {
      const unsigned long long act = ctl->act;
      const unsigned long long odd = ld_agent(&ctl->odd) & act;
      const unsigned long long even = ld_agent(&ctl->even) & act & ~odd;
      const unsigned long long anys = ld_agent(&ctl->any_s), anyt = ld_agent(&ctl->any_t);
      const bool es = d.sides & kSideS, et = d.sides & kSideT;
      unsigned long long alive = act;
      if (es) alive &= anys;
      if (et) alive &= anyt;
      const unsigned long long dead = act & ~alive & ~odd & ~even;
      const int base = ctl->lvl[0] + ctl->lvl[1];
      for (int q = 0; q < 64; ++q) {
        const unsigned long long b = 1ull << q;
        if (odd & b)
          p.dist[wave0 + q] = base + 1;
        else if (even & b)
          p.dist[wave0 + q] = base + 2;
        else if (dead & b)
          p.dist[wave0 + q] = -1;
      }
      ctl->act = act & ~(odd | even | dead);
      ctl->odd = ctl->even = ctl->any_s = ctl->any_t = 0;
      ctl->touched_total += ctl->touched_n;
      ctl->touched_n = 0;
      ctl->steps++;
      ctl->levels_total++;
      for (int x = 0; x < 2; ++x) {
        if (!(d.sides & (1 << x))) continue;
        const int c = ctl->cur[x];
        ctl->lvl[x]++;
        // The consumed parity becomes the next "next": its counters restart.
        // After a bottom-up step its lists are still needed to clear the
        // masks, so kBClearF restarts them.
        if (phase == kBCommit) {
          ctl->fl_n[x][c] = 0;
          ctl->hub[x][c] = 0;
        }
        ctl->fl_mf[x][c] = 0;
        ctl->cur[x] = c ^ 1;
        ctl->fl_mf[x][c ^ 1] = ld_agent(&ctl->fl_mf[x][c ^ 1]);
      }
      if (phase == kBBottomUp) {
        ctl->bu_side = (d.sides & kSideS) ? 0 : 1;
        ctl->phase = kBClearF;  // the consumed masks are cleared first
        return;
      }
      if (!ctl->act) {
        ctl->phase = kBCleanup;
        return;
      }
      batch_decide(ctl, p.mlp, p.V, p.E);
      ctl->phase = ctl->dec.strat == kBottomUp ? kBBottomUp : kBExpand;
      return;
    }
    case kBClearF: {
      const int x = ctl->bu_side;
      ctl->fl_n[x][ctl->cur[x] ^ 1] = 0;
      ctl->hub[x][ctl->cur[x] ^ 1] = 0;
      if (!ctl->act) {
        ctl->phase = kBCleanup;
        return;
      }
      batch_decide(ctl, p.mlp, p.V, p.E);
      ctl->phase = ctl->dec.strat == kBottomUp ? kBBottomUp : kBExpand;
      return;
    }
    case kBCleanup: {
      for (int x = 0; x < 2; ++x) {
        for (int c = 0; c < 2; ++c) {
          ctl->fl_n[x][c] = 0;
          ctl->hub[x][c] = 0;
          ctl->fl_mf[x][c] = 0;
        }
        ctl->cur[x] = 0;
      }
      ctl->dirty_n = 0;
      ctl->touched_n = 0;
      ctl->wave++;
      ctl->phase = ctl->wave < ctl->nwaves ? kBInit : kBIdle;
      return;
    }
    default:
      ctl->phase = kBIdle;
  }
}

template <typename OffT>
__global__ __launch_bounds__(kBlock) void batch_kernel(BatchParams<OffT> p) {
  BatchCtl* ctl = p.ctl;
  const int phase = ctl->phase;
  if (phase == kBIdle) return;
  const uint32_t nb = gridDim.x;
  const int gwave = (int)(blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6));
  const int nwaves = (int)(nb * kWavesPerBlock);
  const int64_t gtid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t nthr = (int64_t)nb * kBlock;
  __shared__ int32_t s_stage[kBLists][kWavesPerBlock * kBStage];
  __shared__ BShared sh;
  const BatchDecision d = ctl->dec;
  const unsigned long long act = ctl->act;
  const int w = (int)(threadIdx.x >> 6);
  Stage s0{&s_stage[0][w * kBStage], 0u}, s1{&s_stage[1][w * kBStage], 0u}, s2{&s_stage[2][w * kBStage], 0u};

  switch (phase) {
    case kBInit: {
      // Seed the wave: one thread per (query, side), block 0 only.
      if (blockIdx.x == 0 && threadIdx.x < 128) {
        const int q = (int)(threadIdx.x & 63), x = (int)(threadIdx.x >> 6);
        const int qi = ctl->wave * 64 + q;
        if (qi < ctl->nq) {
          const int32_t s = p.qs[qi], t = p.qt[qi];
          if (s != t && (uint32_t)s < (uint32_t)p.V && (uint32_t)t < (uint32_t)p.V) {
            const int32_t u = x ? t : s;
            const unsigned long long b = 1ull << q;
            const unsigned long long oldv = atomicOr(&p.rec[4 * (int64_t)u + 2 * x], b);
            if (oldv == 0) {
              const uint32_t k = atomicAdd(&ctl->dirty_n, 1u);
              if (k < (uint32_t)p.V + 256u) p.dirty[k] = u;
            }
            const unsigned long long oldf = atomicOr(&p.F[x][0][u], b);
            if (oldf == 0) {
              const int64_t deg = bdeg(p, u);
              if (deg >= p.hub_threshold) {
                const unsigned long long o =
                    atomicAdd(&ctl->hub[x][0], (1ull << kHubEdgeBits) | (unsigned long long)deg);
                const uint32_t hs = (uint32_t)(o >> kHubEdgeBits);
                if (hs < p.hcap) {
                  p.hl[x][0][hs] = u;
                  p.hb[x][0][hs] = (int64_t)(o & kHubEdgeMask);
                }
              } else {
                const uint32_t k = atomicAdd(&ctl->fl_n[x][0], 1u);
                if (k < (uint32_t)p.V) p.fl[x][0][k] = u;
              }
              atomicAdd(&ctl->fl_mf[x][0], (unsigned long long)deg);
            }
          }
        }
      }
      break;
    }
    case kBExpand: {
      unsigned long long a_cnt = 0, e_cnt = 0;
      const int chunk = ctl->tile > 0 ? ctl->tile : 256;
#pragma unroll
      for (int x = 0; x < 2; ++x)
        if (d.sides & (1 << x)) expand_side(p, s0, x, d.vw, chunk, act, gwave, nwaves, a_cnt, e_cnt);
      uint32_t* counters[1] = {&ctl->touched_n};
      int32_t* lists[1] = {p.touched};
      const uint32_t caps[1] = {(uint32_t)p.V};
      stages_block_flush(s0, s1, s2, 1, counters, lists, caps, &ctl->err, sh);
      unsigned long long vals[2] = {a_cnt, e_cnt};
      const int op[2] = {0, 0};
      unsigned long long* dst[2] = {(unsigned long long*)&ctl->active_vertices,
                                    (unsigned long long*)&ctl->active_edges};
      block_reduce_apply(vals, op, 2, dst, sh);
      break;
    }
    case kBCommit: {
      // hub frontier masks are cleared here: every expand wave is done
      for (int x = 0; x < 2; ++x) {
        if (!(d.sides & (1 << x))) continue;
        const int c = ctl->cur[x];
        uint32_t nh = (uint32_t)(ctl->hub[x][c] >> kHubEdgeBits);
        if (nh > p.hcap) nh = p.hcap;
        for (int64_t i = gtid; i < nh; i += nthr) {
          const int32_t v = p.hl[x][c][i];
          if ((uint32_t)v < (uint32_t)p.V) p.F[x][c][v] = 0ull;
        }
      }
      const int nxt[2] = {ctl->cur[0] ^ 1, ctl->cur[1] ^ 1};
      unsigned long long red[6] = {0, 0, 0, 0, 0, 0};
      uint32_t n = ctl->touched_n;
      if (n > (uint32_t)p.V) n = (uint32_t)p.V;
      for (int64_t base = (int64_t)gwave * kWave; base < n; base += (int64_t)nwaves * kWave) {
        const int64_t i = base + lane_id();
        const bool valid = i < n;
        int32_t u = valid ? p.touched[i] : 0;
        const bool ok = valid && (uint32_t)u < (uint32_t)p.V;
        commit_vertex(p, s0, s1, s2, ok, u, act, nxt, red);
      }
      uint32_t* counters[3] = {&ctl->fl_n[0][nxt[0]], &ctl->fl_n[1][nxt[1]], &ctl->dirty_n};
      int32_t* lists[3] = {p.fl[0][nxt[0]], p.fl[1][nxt[1]], p.dirty};
      const uint32_t caps[3] = {(uint32_t)p.V, (uint32_t)p.V, (uint32_t)p.V + 256u};
      stages_block_flush(s0, s1, s2, 3, counters, lists, caps, &ctl->err, sh);
      const int op[6] = {1, 1, 1, 1, 0, 0};
      unsigned long long* dst[6] = {&ctl->odd, &ctl->even, &ctl->any_s, &ctl->any_t, &ctl->fl_mf[0][nxt[0]],
                                    &ctl->fl_mf[1][nxt[1]]};
      block_reduce_apply(red, op, 6, dst, sh);
      break;
    }
    case kBBottomUp: {
      const int x = (d.sides & kSideS) ? 0 : 1;
      unsigned long long red[6] = {0, 0, 0, 0, 0, 0};
      unsigned long long a_cnt = 0, e_cnt = 0;
      bottom_up(p, s0, s1, s2, x, act, gwave, nwaves, red, a_cnt, e_cnt);
      const int nx0 = ctl->cur[0] ^ 1, nx1 = ctl->cur[1] ^ 1;
      uint32_t* counters[3] = {&ctl->fl_n[0][nx0], &ctl->fl_n[1][nx1], &ctl->dirty_n};
      int32_t* lists[3] = {p.fl[0][nx0], p.fl[1][nx1], p.dirty};
      const uint32_t caps[3] = {(uint32_t)p.V, (uint32_t)p.V, (uint32_t)p.V + 256u};
      stages_block_flush(s0, s1, s2, 3, counters, lists, caps, &ctl->err, sh);
      const int op[8] = {1, 1, 1, 1, 0, 0, 0, 0};
      unsigned long long vals[8] = {red[0], red[1], red[2], red[3], red[4], red[5], a_cnt, e_cnt};
      unsigned long long* dst[8] = {&ctl->odd, &ctl->even, &ctl->any_s, &ctl->any_t,
                                    &ctl->fl_mf[0][ctl->cur[0] ^ 1], &ctl->fl_mf[1][ctl->cur[1] ^ 1],
                                    (unsigned long long*)&ctl->active_vertices,
                                    (unsigned long long*)&ctl->active_edges};
      block_reduce_apply(vals, op, 8, dst, sh);
      break;
    }
    case kBClearF: {
      const int x = ctl->bu_side;
      const int c = ctl->cur[x] ^ 1;  // the parity consumed by the bottom-up step
      const uint32_t n = ctl->fl_n[x][c] < (uint32_t)p.V ? ctl->fl_n[x][c] : (uint32_t)p.V;
      for (int64_t i = gtid; i < n; i += nthr) {
        const int32_t v = p.fl[x][c][i];
        if ((uint32_t)v < (uint32_t)p.V) p.F[x][c][v] = 0ull;
      }
      uint32_t nh = (uint32_t)(ctl->hub[x][c] >> kHubEdgeBits);
      if (nh > p.hcap) nh = p.hcap;
      for (int64_t i = gtid; i < nh; i += nthr) {
        const int32_t v = p.hl[x][c][i];
        if ((uint32_t)v < (uint32_t)p.V) p.F[x][c][v] = 0ull;
      }
      break;
    }
    case kBCleanup: {
      uint32_t nd = ctl->dirty_n;
      if (nd > (uint32_t)p.V + 256u) nd = (uint32_t)p.V + 256u;
      for (int64_t i = gtid; i < nd; i += nthr) {
        const int32_t v = p.dirty[i];
        if ((uint32_t)v < (uint32_t)p.V) {
          ulonglong2* r = reinterpret_cast<ulonglong2*>(p.rec + 4 * (int64_t)v);
          r[0] = make_ulonglong2(0ull, 0ull);
          r[1] = make_ulonglong2(0ull, 0ull);
        }
      }
      for (int x = 0; x < 2; ++x) {
        const int c = ctl->cur[x];
        const uint32_t n = ctl->fl_n[x][c] < (uint32_t)p.V ? ctl->fl_n[x][c] : (uint32_t)p.V;
        for (int64_t i = gtid; i < n; i += nthr) {
          const int32_t v = p.fl[x][c][i];
          if ((uint32_t)v < (uint32_t)p.V) p.F[x][c][v] = 0ull;
        }
        uint32_t nh = (uint32_t)(ctl->hub[x][c] >> kHubEdgeBits);
        if (nh > p.hcap) nh = p.hcap;
        for (int64_t i = gtid; i < nh; i += nthr) {
          const int32_t v = p.hl[x][c][i];
          if ((uint32_t)v < (uint32_t)p.V) p.F[x][c][v] = 0ull;
        }
      }
      break;
    }
    default:
      break;
  }

  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t t = __hip_atomic_fetch_add(&ctl->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    sh.last = t == nb - 1;
  }
  __syncthreads();
  if (!sh.last || threadIdx.x != 0) return;
  batch_finalize(p, phase, d);
  __hip_atomic_store(&ctl->ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename OffT>
hipError_t launch_batch_phase(const BatchParams<OffT>& p, int blocks, hipStream_t s) {
  hipLaunchKernelGGL(batch_kernel<OffT>, dim3(blocks), dim3(kBlock), 0, s, p);
  return hipGetLastError();
}

template <typename OffT>
int batch_occupancy() {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_kernel<OffT>, kBlock, 0) != hipSuccess) return 0;
  return nb;
}

template hipError_t launch_batch_phase<int32_t>(const BatchParams<int32_t>&, int, hipStream_t);
template hipError_t launch_batch_phase<int64_t>(const BatchParams<int64_t>&, int, hipStream_t);
template int batch_occupancy<int32_t>();
template int batch_occupancy<int64_t>();

// ---------------------------------------------------------------------------
// Host driver
// ---------------------------------------------------------------------------
struct BatchWorkspace {
  char* mem = nullptr;
  size_t bytes = 0;
  unsigned long long* rec = nullptr;
  unsigned long long* F[2][2] = {};
  uint32_t* tb = nullptr;
  int32_t* fl[2][2] = {};
  int32_t* hl[2][2] = {};
  int64_t* hb[2][2] = {};
  int32_t* touched = nullptr;
  int32_t* dirty = nullptr;
  BatchCtl* ctl = nullptr;
  int32_t* qs = nullptr;
  int32_t* qt = nullptr;
  int32_t* dist = nullptr;
  int32_t qcap = 0;
  int blocks = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int graph_force = -2;
  int tuned_vw = -1;       // autotuner result (0 = degree-chosen)
  int tuned_tile = -1;
  double est_launches = 12.0;
  int32_t* h_dist = nullptr;  // pinned
  int32_t* h_phase = nullptr; // pinned
};

void batch_free(BatchWorkspace* b) {
  if (!b) return;
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  (void)hipFree(b->mem);
  (void)hipFree(b->qs);
  (void)hipFree(b->qt);
  (void)hipFree(b->dist);
  if (b->h_dist) (void)hipHostFree(b->h_dist);
  if (b->h_phase) (void)hipHostFree(b->h_phase);
  delete b;
}

static fc_status batch_alloc(fc_graph* g) {
  if (g->batch) return FC_OK;
  BatchWorkspace* b = new BatchWorkspace();
  const size_t V = (size_t)g->V;
  const size_t hcap = g->hcap;
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  size_t o = 0;
  const size_t o_rec = o;
  o += al(32 * V);
  const size_t o_F = o;
  o += 4 * al(8 * V);
  const size_t o_tb = o;
  o += al(4 * ((V + 31) / 32));
  const size_t o_fl = o;
  o += 4 * al(4 * V);
  const size_t o_hl = o;
  o += 4 * al(4 * hcap);
  const size_t o_hb = o;
  o += 4 * al(8 * hcap);
  const size_t o_touch = o;
  o += al(4 * V);
  const size_t o_dirty = o;
  o += al(4 * (V + 256));
  const size_t o_ctl = o;
  o += al(sizeof(BatchCtl));
  b->bytes = o;
  hipError_t e = hipMalloc((void**)&b->mem, b->bytes);
  if (e != hipSuccess) {
    delete b;
    return hip_fail(e, "batch workspace allocation");
  }
  e = hipMemset(b->mem, 0, b->bytes);
  if (e != hipSuccess) {
    batch_free(b);
    return hip_fail(e, "batch workspace clear");
  }
  b->rec = (unsigned long long*)(b->mem + o_rec);
  for (int x = 0; x < 2; ++x) {
    for (int c = 0; c < 2; ++c) {
      b->F[x][c] = (unsigned long long*)(b->mem + o_F + (size_t)(x * 2 + c) * al(8 * V));
      b->fl[x][c] = (int32_t*)(b->mem + o_fl + (size_t)(x * 2 + c) * al(4 * V));
      b->hl[x][c] = (int32_t*)(b->mem + o_hl + (size_t)(x * 2 + c) * al(4 * hcap));
      b->hb[x][c] = (int64_t*)(b->mem + o_hb + (size_t)(x * 2 + c) * al(8 * hcap));
    }
  }
  b->tb = (uint32_t*)(b->mem + o_tb);
  b->touched = (int32_t*)(b->mem + o_touch);
  b->dirty = (int32_t*)(b->mem + o_dirty);
  b->ctl = (BatchCtl*)(b->mem + o_ctl);
  int occ = g->off64 ? batch_occupancy<int64_t>() : batch_occupancy<int32_t>();
  if (occ < 1) {
    batch_free(b);
    return fail(FC_ERR_NOGPU, "batch kernel has no occupancy on this device");
  }
  b->blocks = std::min(occ, 8) * g->cus;
  if (hipHostMalloc((void**)&b->h_phase, 4 * sizeof(int32_t), hipHostMallocDefault) != hipSuccess) {
    batch_free(b);
    return fail(FC_ERR_NOMEM, "pinned allocation failed");
  }
  g->batch = b;
  return FC_OK;
}

static fc_status batch_queries(fc_graph* g, int32_t n) {
  BatchWorkspace* b = g->batch;
  if (n <= b->qcap) return FC_OK;
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
  (void)hipFree(b->qs);
  (void)hipFree(b->qt);
  (void)hipFree(b->dist);
  if (b->h_dist) (void)hipHostFree(b->h_dist);
  b->qs = b->qt = b->dist = b->h_dist = nullptr;
  b->qcap = 0;
  const int32_t cap = std::max<int32_t>(n, 4096);
  FC_HIP(hipMalloc((void**)&b->qs, sizeof(int32_t) * (size_t)cap));
  FC_HIP(hipMalloc((void**)&b->qt, sizeof(int32_t) * (size_t)cap));
  FC_HIP(hipMalloc((void**)&b->dist, sizeof(int32_t) * (size_t)cap));
  FC_HIP(hipHostMalloc((void**)&b->h_dist, sizeof(int32_t) * (size_t)cap, hipHostMallocDefault));
  b->qcap = cap;
  return FC_OK;
}

template <typename OffT>
static BatchParams<OffT> batch_params(fc_graph* g, int force_action) {
  BatchWorkspace* b = g->batch;
  BatchParams<OffT> p;
  std::memset(&p, 0, sizeof(p));
  p.off = (const OffT*)g->d_off;
  p.col = g->d_col;
  p.V = (int32_t)g->V;
  p.hub_threshold = g->hub_threshold;
  p.E = g->E;
  p.rec = b->rec;
  for (int x = 0; x < 2; ++x) {
    for (int c = 0; c < 2; ++c) {
      p.F[x][c] = b->F[x][c];
      p.fl[x][c] = b->fl[x][c];
      p.hl[x][c] = b->hl[x][c];
      p.hb[x][c] = b->hb[x][c];
    }
  }
  p.hcap = g->hcap;
  p.tb = b->tb;
  p.touched = b->touched;
  p.dirty = b->dirty;
  p.ctl = b->ctl;
  p.qs = b->qs;
  p.qt = b->qt;
  p.dist = b->dist;
  p.mlp = g->d_mlp;
  p.force_action = force_action;
  return p;
}

static hipError_t batch_launch_raw(fc_graph* g, int force_action) {
  if (g->off64)
    return launch_batch_phase<int64_t>(batch_params<int64_t>(g, force_action), g->batch->blocks, g->stream);
  return launch_batch_phase<int32_t>(batch_params<int32_t>(g, force_action), g->batch->blocks, g->stream);
}

fc_status batch_launch_once(fc_graph* g, int force_action) {
  hipError_t e = batch_launch_raw(g, force_action);
  if (e != hipSuccess) return hip_fail(e, "batch kernel launch");
  return FC_OK;
}

static fc_status batch_graph(fc_graph* g) {
  BatchWorkspace* b = g->batch;
  if (b->gexec) return FC_OK;
  FC_HIP(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
  hipError_t le = hipSuccess;
  for (int i = 0; i < kGraphLaunches && le == hipSuccess; ++i) le = batch_launch_raw(g, -1);
  hipGraph_t gr = nullptr;
  hipError_t ce = hipStreamEndCapture(g->stream, &gr);
  if (le != hipSuccess) return hip_fail(le, "batch kernel capture");
  if (ce != hipSuccess) return hip_fail(ce, "hipStreamEndCapture");
  b->graph = gr;
  FC_HIP(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
  return FC_OK;
}

// Prepare the control block for a batch of n queries (already on the device).
fc_status batch_begin(fc_graph* g, int32_t n, int policy, int vw, int tile) {
  BatchWorkspace* b = g->batch;
  BatchCtl init;
  std::memset(&init, 0, sizeof(init));
  init.phase = kBInit;
  init.wave = 0;
  init.nwaves = (n + 63) / 64;
  init.nq = n;
  init.policy = policy;
  init.vw_fixed = vw;
  init.tile = tile;
  FC_HIP(hipMemcpyAsync(b->ctl, &init, sizeof(init), hipMemcpyHostToDevice, g->stream));
  return FC_OK;
}

// Run the phase machine until idle. use_graph: replay captured launches.
static fc_status batch_drive(fc_graph* g, int32_t n, bool use_graph, float* ms) {
  BatchWorkspace* b = g->batch;
  const int64_t nw = (n + 63) / 64;
  int64_t expected = (int64_t)std::ceil((double)nw * b->est_launches) + 2;
  int64_t launched = 0;
  FC_HIP(hipEventRecord(g->ev0, g->stream));
  for (int round = 0;; ++round) {
    int64_t want = std::max<int64_t>(expected - launched, 1);
    if (!use_graph || want < kGraphLaunches) {
      for (int64_t i = 0; i < want; ++i) {
        hipError_t e = batch_launch_raw(g, -1);
        if (e != hipSuccess) return hip_fail(e, "batch kernel launch");
      }
      launched += want;
    } else {
      const int64_t ng = (want + kGraphLaunches - 1) / kGraphLaunches;
      for (int64_t i = 0; i < ng; ++i) FC_HIP(hipGraphLaunch(b->gexec, g->stream));
      launched += ng * kGraphLaunches;
    }
    FC_HIP(hipEventRecord(g->ev1, g->stream));
    FC_HIP(hipMemcpyAsync(b->h_phase, b->ctl, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, g->stream));
    FC_HIP(hipStreamSynchronize(g->stream));
    if (b->h_phase[0] == kBIdle) break;
    if (launched > nw * (4 * g->V + 16) + 64)
      return fail(FC_ERR_CAPACITY, "batch phase machine did not finish after %lld launches", (long long)launched);
    const int64_t remaining = std::max<int64_t>(nw - b->h_phase[1], 1);
    expected = launched + (int64_t)std::ceil((double)remaining * (b->est_launches + 4.0) * (1.0 + round)) + 2;
  }
  FC_HIP(hipEventElapsedTime(ms, g->ev0, g->ev1));
  b->est_launches = 0.5 * b->est_launches + 0.5 * ((double)launched / (double)nw);
  return FC_OK;
}

static fc_status batch_run(fc_graph* g, const int32_t* s, const int32_t* t, int32_t n, int policy, int vw, int tile,
                           bool use_graph, float* ms, BatchCtl* ctl_out) {
  BatchWorkspace* b = g->batch;
  fc_status st = batch_queries(g, n);
  if (st != FC_OK) return st;
  if (use_graph) {
    st = batch_graph(g);
    if (st != FC_OK) return st;
  }
  FC_HIP(hipMemcpyAsync(b->qs, s, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, g->stream));
  FC_HIP(hipMemcpyAsync(b->qt, t, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, g->stream));
  st = batch_begin(g, n, policy, vw, tile);
  if (st != FC_OK) return st;
  FC_HIP(hipStreamSynchronize(g->stream));
  st = batch_drive(g, n, use_graph, ms);
  if (st != FC_OK) return st;
  FC_HIP(hipMemcpyAsync(b->h_dist, b->dist, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost, g->stream));
  if (ctl_out) FC_HIP(hipMemcpyAsync(ctl_out, b->ctl, sizeof(BatchCtl), hipMemcpyDeviceToHost, g->stream));
  FC_HIP(hipStreamSynchronize(g->stream));
  if (ctl_out && ctl_out->err) return fail(FC_ERR_CAPACITY, "batch kernel: list guard tripped (err=%u)", ctl_out->err);
  return FC_OK;
}

// Autotuner: time the first waves of this batch under each (width, tile)
// candidate and keep the fastest for this graph.
static fc_status batch_autotune(fc_graph* g, const int32_t* s, const int32_t* t, int32_t n, int policy) {
  BatchWorkspace* b = g->batch;
  const int32_t m = std::min<int32_t>(n, 128);
  const int vws[] = {0, 8, 32, 64};
  const int tiles[] = {256, 1024};
  float best = 1e30f;
  int bvw = 0, btile = 256;
  for (int vw : vws)
    for (int tile : tiles) {
      float ms = 0;
      BatchCtl c;
      fc_status st = batch_run(g, s, t, m, policy, vw, tile, false, &ms, &c);
      if (st != FC_OK) return st;
      if (ms < best) {
        best = ms;
        bvw = vw;
        btile = tile;
      }
    }
  b->tuned_vw = bvw;
  b->tuned_tile = btile;
  return FC_OK;
}

}  // namespace fcd

using fcd::BatchCtl;

extern "C" fc_status fc_stcon_batch(fc_graph* g, const int32_t* s, const int32_t* t, int32_t n, int32_t* connected,
                                    int32_t* dist, const fc_batch_opts* opts, fc_batch_stats* stats) {
  fc::clear_error();
  if (!g || (n > 0 && (!s || !t))) return fc::fail(FC_ERR_INVALID, "fc_stcon_batch: null argument");
  if (n < 0) return fc::fail(FC_ERR_INVALID, "fc_stcon_batch: negative count");
  for (int32_t i = 0; i < n; ++i)
    if (s[i] < 0 || s[i] >= g->V || t[i] < 0 || t[i] >= g->V)
      return fc::fail(FC_ERR_INVALID, "fc_stcon_batch: query %d vertex out of range [0,%lld)", i, (long long)g->V);
  fc_batch_opts o;
  std::memset(&o, 0, sizeof(o));
  o.policy = FC_POLICY_HEURISTIC;
  o.use_graph = 1;
  o.words = 1;
  if (opts) o = *opts;
  if (o.words != 0 && o.words != 1) return fc::fail(FC_ERR_UNSUPPORTED, "fc_stcon_batch: only words=1 is built");
  if (o.policy < 0 || o.policy > 2) return fc::fail(FC_ERR_INVALID, "fc_stcon_batch: unknown policy %d", o.policy);
  if (o.policy == FC_POLICY_LEARNED && !g->mlp_set)
    return fc::fail(FC_ERR_INVALID, "fc_stcon_batch: learned policy requested but no scheduler parameters are set");
  if (o.vw != 0 && o.vw != 4 && o.vw != 8 && o.vw != 16 && o.vw != 32 && o.vw != 64)
    return fc::fail(FC_ERR_INVALID, "fc_stcon_batch: vw must be 0, 4, 8, 16, 32 or 64");
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (n == 0) return FC_OK;
  fc_status st = fcd::batch_alloc(g);
  if (st != FC_OK) return st;
  g->levels_valid = g->parents_valid = false;
  fcd::BatchWorkspace* b = g->batch;
  int vw = o.vw, tile = o.tile;
  if (vw == 0 && tile == 0) {
    if (b->tuned_vw < 0) {
      st = fcd::batch_autotune(g, s, t, n, o.policy);
      if (st != FC_OK) return st;
    }
    vw = b->tuned_vw;
    tile = b->tuned_tile;
  }
  if (tile == 0) tile = 256;
  float ms = 0;
  BatchCtl c;
  st = fcd::batch_run(g, s, t, n, o.policy, vw, tile, o.use_graph != 0, &ms, &c);
  if (st != FC_OK) return st;
  for (int32_t i = 0; i < n; ++i) {
    const int32_t d = b->h_dist[i];
    if (dist) dist[i] = d;
    if (connected) connected[i] = d >= 0 ? 1 : 0;
  }
  if (stats) {
    stats->time_ms = ms;
    stats->waves = (n + 63) / 64;
    stats->levels = (int32_t)c.levels_total;
    stats->active_vertices = c.active_vertices;
    stats->active_edges = c.active_edges;
    stats->touched_vertices = c.touched_total;
    stats->vw = vw;
    stats->tile = tile;
  }
  return FC_OK;
}
