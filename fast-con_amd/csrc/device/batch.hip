// Batched s–t connectivity: waves of 64 queries answered by one bit-parallel
// bidirectional visit (one 64-bit mask per vertex and side), so that the
// adjacency of hub vertices, which almost every R-MAT query touches, is read
// once per wave instead of once per query.
//
// Per vertex u and side x (0 = s side, 1 = t side):
//   rec[2Vx+2u]   queries whose side-x ball contains u
//   rec[2Vx+2u+1] masks ORed in during a top-down expansion, consumed by commit
//   F[x][p][u]    queries for which u is in side x's current frontier
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
// Only the masks a wave touched are cleared afterwards (dirty bitmap), so a
// wave costs nothing proportional to V beyond the bottom-up steps.
#include "batch.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "engine.hpp"
#include "hubstage.hpp"
#include "mlp.hpp"

namespace fcd {

using fc::fail;

constexpr uint32_t kBStage = 512;  // staged entries per wave and list
// The control block's state: the words before the phase log, which the
// finalising wave copies to LDS.
constexpr uint32_t kCtlHotWords = offsetof(BatchCtl, plog) / sizeof(unsigned long long);
static_assert(offsetof(BatchCtl, plog) % sizeof(unsigned long long) == 0, "state ends on a word boundary");
static_assert(kCtlHotWords <= 2 * kWave, "two loads per lane at most");
// Margin of the learned policy when the caller gives none: the value model
// selection chose on the validation graphs for the shipped parameters
// (fastcon/sched_report.json), so C callers, the CLIs and the Python binding
// all get the same behaviour.
constexpr float kDefaultLearnMargin = 1.0f;
constexpr int kMaxLanes = 4;          // concurrent wave machines per batch
constexpr int kDefaultLanes = 3;
// A filter launch pays when the expected number of dropped list entries or
// frontier edges reaches these (defaults; FASTCON_FILTER_MIN_VERTICES and
// FASTCON_FILTER_MIN_EDGES override them).
constexpr uint64_t kFilterMinVertices = 1u << 15;
constexpr uint64_t kFilterMinEdges = 1u << 20;

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
  uint32_t cnt[4][kWavesPerBlock];              // staged entries: lists s, t, hubs s, t
  unsigned long long hedge[2][kWavesPerBlock];  // staged hub edges per side
  uint32_t base[2];
  unsigned long long hbase[2];
  unsigned long long red[kRCount][kWavesPerBlock];
  int last;
};

template <typename OffT>
__device__ __forceinline__ int64_t bdeg(const BatchParams<OffT>& p, int32_t u) {
  return (int64_t)(p.off[u + 1] - p.off[u]);
}

// Block reduction of the kRCount launch values (op: OR for the first four,
// sum for the rest) into this block's slot; lanes 0..kRCount-1 of wave 0
// issue one atomic each, skipping zeros.
__device__ __forceinline__ void block_reduce_slots(const unsigned long long* vals, unsigned long long* red,
                                                   BShared& sh) {
  const int w = (int)(threadIdx.x >> 6);
#pragma unroll
  for (int k = 0; k < kRCount; ++k) {
    const unsigned long long v = k < kRMfS ? (unsigned long long)wave_reduce_full_nz_u64<kWaveOr>(vals[k])
                                           : (unsigned long long)wave_reduce_full_nz_u64<kWaveSum>(vals[k]);
    if (lane_id() == 0) sh.red[k][w] = v;
  }
  __syncthreads();
  if (threadIdx.x < kRCount) {
    const int k = (int)threadIdx.x;
    unsigned long long a = 0;
    for (int i = 0; i < kWavesPerBlock; ++i) a = k < kRMfS ? (a | sh.red[k][i]) : (a + sh.red[k][i]);
    unsigned long long* dst = red + (size_t)(blockIdx.x % kRedSlots) * kRedWords + k;
    if (a) {
      if (k < kRMfS)
        atomicOr(dst, a);
      else
        atomicAdd(dst, a);
    }
  }
  __syncthreads();
}

struct RedTot {
  unsigned long long v[kRCount];
};

// Per-wave append state of the commit, bottom-up and filter phases: the next
// frontier lists of both sides (normal vertices and hubs).
struct Appends {
  Stage s[2];
  HubStage h[2];
};

// Put a newly reached vertex u of side x into the next frontier: the normal
// list or, by degree, the hub list (whose edge bases the hub stage assigns).
template <typename OffT>
__device__ __forceinline__ void push_frontier(const BatchParams<OffT>& p, Appends& ap, int x, int nxt, bool want,
                                              int32_t u, int64_t deg) {
  const bool hub = want && deg >= p.hub_threshold;
  stage_push(ap.s[x], want && !hub, u, &p.ctl->fl_n[x][nxt], p.fl[x][nxt], (uint32_t)p.V, &p.ctl->err);
  hub_push(ap.h[x], hub, u, deg, &p.ctl->hub[x][nxt], p.hl[x][nxt], p.hb[x][nxt], p.hcap, &p.ctl->err);
}

// Block-collective publication of the staged appends (all threads call,
// uniform control flow): one barrier to publish every wave's counts, the
// four list reservations issued in parallel by threads 0..3, then each wave
// writes its entries.
template <typename OffT>
__device__ __forceinline__ void appends_block_flush(const BatchParams<OffT>& p, Appends& ap, const int* nxt,
                                                    BShared& sh) {
  BatchCtl* ctl = p.ctl;
  const int w = (int)(threadIdx.x >> 6);
  const int lane = lane_id();
  unsigned long long he[2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
    he[x] = wave_reduce_full_nz_u64<kWaveSum>((uint32_t)lane < ap.h[x].cnt ? (unsigned long long)ap.h[x].d[lane] : 0ull);
  if (lane == 0) {
    sh.cnt[0][w] = ap.s[0].cnt;
    sh.cnt[1][w] = ap.s[1].cnt;
    sh.cnt[2][w] = ap.h[0].cnt;
    sh.cnt[3][w] = ap.h[1].cnt;
    sh.hedge[0][w] = he[0];
    sh.hedge[1][w] = he[1];
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    const int l = (int)threadIdx.x;
    uint32_t c = 0;
    unsigned long long e = 0;
    for (int i = 0; i < kWavesPerBlock; ++i) {
      const uint32_t ci = sh.cnt[l][i];
      sh.cnt[l][i] = c;
      c += ci;
      if (l >= 2) {
        const unsigned long long ei = sh.hedge[l - 2][i];
        sh.hedge[l - 2][i] = e;
        e += ei;
      }
    }
    if (l < 2)
      sh.base[l] = c ? atomicAdd(&ctl->fl_n[l][nxt[l]], c) : 0u;
    else
      sh.hbase[l - 2] = c ? atomicAdd(&ctl->hub[l - 2][nxt[l - 2]], ((unsigned long long)c << kHubEdgeBits) | e) : 0ull;
  }
  __syncthreads();
#pragma unroll
  for (int x = 0; x < 2; ++x) {
    Stage& st = ap.s[x];
    const uint32_t base = sh.base[x] + sh.cnt[x][w];
    for (uint32_t i = (uint32_t)lane; i < st.cnt; i += kWave) {
      const uint32_t slot = base + i;
      if (slot < (uint32_t)p.V)
        p.fl[x][nxt[x]][slot] = st.buf[i];
      else
        atomicOr(&ctl->err, 1u);
    }
    st.cnt = 0;
    HubStage& h = ap.h[x];
    if (h.cnt) {
      const unsigned long long hb = sh.hbase[x];
      hub_write(h, (uint32_t)(hb >> kHubEdgeBits) + sh.cnt[2 + x][w], (hb & kHubEdgeMask) + sh.hedge[x][w],
                p.hl[x][nxt[x]], p.hb[x][nxt[x]], p.hcap, &ctl->err);
    }
    h.cnt = 0;
  }
}

// Top-down probe of edge (v -> u) for side x with frontier mask fm, split in
// a load and an update so that a lane's loads for several edges are in
// flight together. The update ORs the new bits into u's N word with an
// atomic that returns nothing and, when the loaded N word was zero, marks
// u's group of 32 vertices with a plain byte store (every writer stores the
// same value, so the race is benign). Every vertex whose N word becomes
// non-zero is in a marked group: the first atomic on it is issued by a lane
// whose earlier load saw zero.
// Side x's {visited, N} pair of vertex u: the records are side-major, so a
// one-side top-down probe reads 16 bytes of an array in which every record
// belongs to that side.
template <typename OffT>
__device__ __forceinline__ unsigned long long* rec_of(const BatchParams<OffT>& p, int x, int64_t u) {
  return p.rec + 2 * (int64_t)p.V * x + 2 * u;
}

template <typename OffT>
__device__ __forceinline__ ulonglong2 probe_load(const BatchParams<OffT>& p, int x, unsigned long long fm, int32_t u,
                                                 bool act) {
  ulonglong2 vn = make_ulonglong2(~0ull, 0ull);
  if (act && fm) vn = *reinterpret_cast<const ulonglong2*>(rec_of(p, x, u));  // {visited, next}
  return vn;
}

template <typename OffT>
__device__ __forceinline__ void probe_apply(const BatchParams<OffT>& p, int x, unsigned long long fm, int32_t u,
                                            ulonglong2 vn) {
  const unsigned long long m = fm & ~vn.x;
  if (m & ~vn.y) {
    __hip_atomic_fetch_or(rec_of(p, x, u) + 1, m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (vn.y == 0) p.tm[u >> 5] = 1;
  }
}

template <typename OffT, int VW>
__device__ __forceinline__ void expand_normal(const BatchParams<OffT>& p, int x, int par, uint32_t n,
                                              unsigned long long act, int gwave, int nwaves,
                                              unsigned long long& a_cnt, unsigned long long& e_cnt) {
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
      const ulonglong2 r0 = probe_load(p, x, fm, u0, a0);
      const ulonglong2 r1 = probe_load(p, x, fm, u1, a1);
      if (a0) probe_apply(p, x, fm, u0, r0);
      if (a1) probe_apply(p, x, fm, u1, r1);
      b += 2 * VW;
    }
  }
}

// Hub frontier vertices are expanded edge-parallel: the concatenated
// adjacency of the hub list is cut into chunks of `chunk` edges (the tile),
// one chunk per wave at a time, four edges per lane and pass. When the hub
// edges are too few for every wave to get a whole tile, the chunk is halved
// (down to 64 edges, staying a multiple of 64) so that the work spreads over
// more waves: a wave's chain of dependent loads per pass, not the edge
// count, then sets the phase's time (the roots' expansion at the start of a
// wave has a few thousand hub edges).
template <typename OffT>
__device__ __forceinline__ void expand_hubs(const BatchParams<OffT>& p, int x, int par, unsigned long long hubpk,
                                            unsigned long long act, int tile, int gwave, int nwaves,
                                            unsigned long long& a_cnt, unsigned long long& e_cnt) {
  constexpr int R = 4;
  uint32_t nh = (uint32_t)(hubpk >> kHubEdgeBits);
  if (nh > p.hcap) nh = p.hcap;
  const uint64_t H = hubpk & kHubEdgeMask;
  if (nh == 0 || H == 0) return;
  int chunk = tile;
  while (chunk > kWave && (uint64_t)chunk * (uint64_t)nwaves > H) {
    const int half = (chunk >> 1) & ~(kWave - 1);
    chunk = half > kWave ? half : kWave;
  }
  const int lane = lane_id();
  const int32_t* hl = p.hl[x][par];
  const int64_t* hb = p.hb[x][par];
  const unsigned long long* F = p.F[x][par];
  for (uint64_t cb = (uint64_t)gwave * (uint64_t)chunk; cb < H; cb += (uint64_t)nwaves * (uint64_t)chunk) {
    // 64-ary search: the last hub whose edge base is <= cb
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
    uint32_t s = lo;
    for (int k = 0; k < chunk; k += R * kWave) {
      OffT pos[R];
      unsigned long long fm[R];
      bool ae[R];
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const uint64_t g = cb + (uint64_t)(k + j * kWave + lane);
        ae[j] = k + j * kWave < chunk && g < H;
        fm[j] = 0;
        pos[j] = 0;
        if (ae[j]) {
          while (s + 1 < nh && (uint64_t)hb[s + 1] <= g) ++s;
          const int32_t v = hl[s];
          const uint64_t b0 = (uint64_t)hb[s];
          bool good = (uint32_t)v < (uint32_t)p.V && b0 <= g;
          if (good) {
            pos[j] = p.off[v] + (OffT)(g - b0);
            good = pos[j] < p.off[v + 1];
          }
          if (good) {
            fm[j] = F[v] & act;
            if (fm[j]) {
              e_cnt += 1;
              if (b0 == g) a_cnt += 1;  // first edge of a hub with live queries
            }
          } else {
            atomicOr(&p.ctl->err, 2u);
            ae[j] = false;
          }
        }
      }
      int32_t u[R];
#pragma unroll
      for (int j = 0; j < R; ++j) u[j] = (ae[j] && fm[j]) ? p.col[pos[j]] : 0;
      ulonglong2 r[R];
#pragma unroll
      for (int j = 0; j < R; ++j) r[j] = probe_load(p, x, fm[j], u[j], ae[j]);
#pragma unroll
      for (int j = 0; j < R; ++j)
        if (ae[j] && fm[j]) probe_apply(p, x, fm[j], u[j], r[j]);
    }
  }
}

template <typename OffT>
__device__ __forceinline__ void expand_side(const BatchParams<OffT>& p, int x, int vw, int chunk,
                                            unsigned long long act, int gwave, int nwaves,
                                            unsigned long long& a_cnt, unsigned long long& e_cnt) {
  const BatchCtl* ctl = p.ctl;
  const int par = ctl->cur[x];
  const uint32_t n = ctl->fl_n[x][par];
  switch (vw) {
    case 4: expand_normal<OffT, 4>(p, x, par, n, act, gwave, nwaves, a_cnt, e_cnt); break;
    case 8: expand_normal<OffT, 8>(p, x, par, n, act, gwave, nwaves, a_cnt, e_cnt); break;
    case 16: expand_normal<OffT, 16>(p, x, par, n, act, gwave, nwaves, a_cnt, e_cnt); break;
    case 32: expand_normal<OffT, 32>(p, x, par, n, act, gwave, nwaves, a_cnt, e_cnt); break;
    default: expand_normal<OffT, 64>(p, x, par, n, act, gwave, nwaves, a_cnt, e_cnt); break;
  }
  expand_hubs(p, x, par, ctl->hub[x][par], act, chunk, gwave, nwaves, a_cnt, e_cnt);
}

// Groups of 32 vertices are dealt to the waves of a launch round-robin:
// wave w owns groups w, w + nwaves, w + 2 nwaves, ... With vertices
// numbered by decreasing degree the touched groups crowd at low ids, and
// this spreads them over every wave. A wave reads the marks of 64 of its
// groups at a time; slot k of such a read is group (jb + k) * nwaves + w.
struct GroupDeal {
  int64_t jb;
  int nwaves, gwave;
  __device__ __forceinline__ int64_t operator()(int k) const { return (jb + k) * (int64_t)nwaves + gwave; }
};

// Number of groups owned by wave gwave (groups = ceil(V / 32) < 2^27, so
// a 32-bit division does).
__device__ __forceinline__ int64_t owned_groups(int64_t groups, int gwave, int nwaves) {
  return groups > gwave ? (int64_t)(((uint32_t)groups - (uint32_t)gwave + (uint32_t)nwaves - 1u) / (uint32_t)nwaves)
                        : 0;
}

// Mark dirty vertices (visited words became non-zero this wave) in the dirty
// bitmap, for a pass whose lanes 0..31 / 32..63 hold the 32 vertices of
// groups g0 / g1 (each group owned by this wave in this launch). dbw is the
// group's dirty word as loaded with the pass's records, so that the update
// waits for no load of its own.
template <typename OffT>
__device__ __forceinline__ void mark_dirty(const BatchParams<OffT>& p, bool dirty, int64_t g0, int64_t g1,
                                           uint32_t dbw) {
  const uint64_t dm = ballot(dirty);
  const int lane = lane_id();
  if (lane == 0 && (uint32_t)dm) p.db[g0] = dbw | (uint32_t)dm;
  if (lane == 32 && (dm >> 32) && g1 >= 0) p.db[g1] = dbw | (uint32_t)(dm >> 32);
}

// Commit: every vertex of a touched group folds its N words into its visited
// words, and newly reached vertices join the next frontiers. The wave reads
// the marks of 64 of its groups (dealt round-robin), then handles the marked
// groups two at a time (lanes 0..31 and 32..63, each half-wave reading and
// writing one group's records in address order); the records and
// offsets of the next pair are loaded before the current pair is processed,
// so each pass costs about one memory round trip. Each group belongs to one
// wave of the launch, which therefore updates its dirty-bitmap word without
// an atomic.
template <typename OffT>
struct CommitLoad {
  ulonglong2 a, b;  // {visS, NS}, {visT, NT}
  OffT o0, o1;
  int32_t u;
  uint32_t dbw;     // the dirty-bitmap word of the lane's group
  bool valid;
};

template <typename OffT>
__device__ __forceinline__ CommitLoad<OffT> commit_load(const BatchParams<OffT>& p, const GroupDeal& gd, int k0,
                                                        int k1) {
  const int lane = lane_id();
  const int mine = lane < 32 ? k0 : k1;
  CommitLoad<OffT> c;
  const int64_t u = gd(mine < 0 ? 0 : mine) * 32 + (lane & 31);
  c.u = (int32_t)u;
  c.valid = mine >= 0 && u < p.V;
  c.a = c.b = make_ulonglong2(0ull, 0ull);
  c.o0 = c.o1 = 0;
  c.dbw = mine >= 0 ? p.db[gd(mine)] : 0u;
  if (c.valid) {
    c.a = *reinterpret_cast<const ulonglong2*>(rec_of(p, 0, u));
    c.b = *reinterpret_cast<const ulonglong2*>(rec_of(p, 1, u));
    c.o0 = p.off[u];
    c.o1 = p.off[u + 1];
  }
  return c;
}

__device__ __forceinline__ bool take_pair(uint64_t& nz, int& k0, int& k1) {
  if (!nz) return false;
  k0 = __builtin_ctzll(nz);
  nz &= nz - 1;
  k1 = -1;
  if (nz) {
    k1 = __builtin_ctzll(nz);
    nz &= nz - 1;
  }
  return true;
}

template <typename OffT>
__device__ __forceinline__ void commit_pass(const BatchParams<OffT>& p, Appends& ap, unsigned long long act,
                                            const int* nxt, int gwave, int nwaves, unsigned long long* red) {
  const int lane = lane_id();
  const int64_t per = owned_groups(((int64_t)p.V + 31) / 32, gwave, nwaves);
  for (int64_t jb = 0; jb < per; jb += kWave) {
    const GroupDeal gd{jb, nwaves, gwave};
    const int64_t gi = gd(lane);
    const bool marked = jb + lane < per && p.tm[gi] != 0;
    if (marked) p.tm[gi] = 0;
    uint64_t nz = ballot(marked);
    int k0, k1;
    if (!take_pair(nz, k0, k1)) continue;
    CommitLoad<OffT> cur = commit_load(p, gd, k0, k1);
    for (;;) {
      int n0 = -1, n1 = -1;
      const bool more = take_pair(nz, n0, n1);
      CommitLoad<OffT> nx;
      if (more) nx = commit_load(p, gd, n0, n1);
      unsigned long long nS = 0, nT = 0, vS = 0, vT = 0;
      int64_t deg = 0;
      const int32_t u = cur.u;
      if (cur.valid) {
        vS = cur.a.x;
        vT = cur.b.x;
        nS = cur.a.y & ~vS & act;
        nT = cur.b.y & ~vT & act;
        red[kROdd] |= (nS & vT) | (nT & vS);  // odd meets
        red[kREven] |= nS & nT;                // even meets
        red[kRAnyS] |= nS;
        red[kRAnyT] |= nT;
        if (cur.a.y | cur.b.y) red[kRTouched] += 1;
        if (cur.a.y) *reinterpret_cast<ulonglong2*>(rec_of(p, 0, u)) = make_ulonglong2(vS | nS, 0ull);
        if (cur.b.y) *reinterpret_cast<ulonglong2*>(rec_of(p, 1, u)) = make_ulonglong2(vT | nT, 0ull);
        if (nS) p.F[0][nxt[0]][u] = nS;
        if (nT) p.F[1][nxt[1]][u] = nT;
        deg = (int64_t)(cur.o1 - cur.o0);
        if (nS) red[kRMfS] += (unsigned long long)deg;
        if (nT) red[kRMfT] += (unsigned long long)deg;
        red[kRVisS] += (unsigned long long)__popcll(nS);
        red[kRVisT] += (unsigned long long)__popcll(nT);
      }
      push_frontier(p, ap, 0, nxt[0], nS != 0, u, deg);
      push_frontier(p, ap, 1, nxt[1], nT != 0, u, deg);
      mark_dirty(p, ((vS | vT) == 0) && ((nS | nT) != 0), gd(k0), k1 >= 0 ? gd(k1) : (int64_t)-1, cur.dbw);
      if (!more) break;
      cur = nx;
      k0 = n0;
      k1 = n1;
    }
  }
}

// Bottom-up step for the sides in SIDES (bit x = side x): every vertex still
// needed by some live query ORs the current frontier masks of its
// neighbours, stopping once it has every bit it needs. Both sides share one
// scan of the adjacency list.
template <typename OffT, int SIDES>
__device__ __forceinline__ void bottom_up(const BatchParams<OffT>& p, Appends& ap, unsigned long long act,
                                          int gwave, int nwaves, unsigned long long* red) {
  const BatchCtl* ctl = p.ctl;
  constexpr bool DS = (SIDES & 1) != 0, DT = (SIDES & 2) != 0;
  const int nxt[2] = {ctl->cur[0] ^ 1, ctl->cur[1] ^ 1};
  const unsigned long long* FS = p.F[0][ctl->cur[0]];
  const unsigned long long* FT = p.F[1][ctl->cur[1]];
  const int lane = lane_id();
  const int64_t nthr = (int64_t)nwaves * kWave;
  // isolated vertices (the tail of a degree-ordered numbering) gain no bit
  for (int64_t base = (int64_t)gwave * kWave; base < p.v_active; base += nthr) {
    const int32_t u = (int32_t)(base + lane);
    unsigned long long accS = 0, accT = 0, needS = 0, needT = 0, vS = 0, vT = 0;
    OffT b = 0, e = 0;
    // this lane's group's dirty word (lanes 0..31: base / 32, 32..63: the next)
    const uint32_t dbw = base + (lane & 32) < p.V ? p.db[(base >> 5) + (lane >> 5)] : 0u;
    if (u < p.V) {
      vS = rec_of(p, 0, u)[0];
      vT = rec_of(p, 1, u)[0];
      if (DS) needS = act & ~vS;
      if (DT) needT = act & ~vT;
      if (needS | needT) {
        b = p.off[u];
        e = p.off[u + 1];
      }
    }
    if (b < e) {
      red[kRActV] += 1;
      for (OffT j = b; j < e; j += 4) {
        int32_t c[4];
        c[0] = p.col[j];
        c[1] = j + 1 < e ? p.col[j + 1] : -1;
        c[2] = j + 2 < e ? p.col[j + 2] : -1;
        c[3] = j + 3 < e ? p.col[j + 3] : -1;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          if (c[k] < 0) continue;
          if (DS) accS |= FS[c[k]];
          if (DT) accT |= FT[c[k]];
        }
        red[kRActE] += (unsigned long long)((j + 4 <= e) ? 4 : (e - j));
        if ((accS & needS) == needS && (accT & needT) == needT) break;
      }
    }
    const unsigned long long nS = accS & needS, nT = accT & needT;
    int64_t deg = 0;
    if (nS | nT) {
      if (nS) rec_of(p, 0, u)[0] = vS | nS;
      if (nT) rec_of(p, 1, u)[0] = vT | nT;
      if (nS) p.F[0][nxt[0]][u] = nS;
      if (nT) p.F[1][nxt[1]][u] = nT;
      red[kROdd] |= (nS & vT) | (nT & vS);
      red[kREven] |= nS & nT;
      red[kRAnyS] |= nS;
      red[kRAnyT] |= nT;
      deg = (int64_t)(e - b);
      if (nS) red[kRMfS] += (unsigned long long)deg;
      if (nT) red[kRMfT] += (unsigned long long)deg;
      red[kRVisS] += (unsigned long long)__popcll(nS);
      red[kRVisT] += (unsigned long long)__popcll(nT);
    }
    if (DS) push_frontier(p, ap, 0, nxt[0], nS != 0, u, deg);
    if (DT) push_frontier(p, ap, 1, nxt[1], nT != 0, u, deg);
    // vertices base .. base+63 are this wave's alone in this launch
    mark_dirty(p, (nS | nT) != 0 && (vS | vT) == 0, base >> 5, (base >> 5) + 1, dbw);
  }
}

// Filter: the current frontier lists of side x still hold vertices of
// queries resolved since they were built. Keep each vertex's mask restricted
// to the live queries in the other parity, clear the old mask, and rebuild
// the lists (and the hub edge bases) from the survivors only.
template <typename OffT, int x>
__device__ __forceinline__ void filter_side(const BatchParams<OffT>& p, Appends& ap, unsigned long long act,
                                            int gwave, int nwaves, unsigned long long* red) {
  const BatchCtl* ctl = p.ctl;
  const int c = ctl->cur[x], nx = c ^ 1;
  const int lane = lane_id();
  unsigned long long* Fc = p.F[x][c];
  unsigned long long* Fn = p.F[x][nx];
  uint32_t n = ctl->fl_n[x][c];
  if (n > (uint32_t)p.V) n = (uint32_t)p.V;
  uint32_t nh = (uint32_t)(ctl->hub[x][c] >> kHubEdgeBits);
  if (nh > p.hcap) nh = p.hcap;
  const uint32_t tot = n + nh;
  for (uint32_t base = (uint32_t)gwave * kWave; base < tot; base += (uint32_t)nwaves * kWave) {
    const uint32_t i = base + (uint32_t)lane;
    int32_t v = -1;
    if (i < tot) v = i < n ? p.fl[x][c][i] : p.hl[x][c][i - n];
    unsigned long long keep = 0;
    int64_t deg = 0;
    if ((uint32_t)v < (uint32_t)p.V) {
      const unsigned long long fm = Fc[v];
      Fc[v] = 0ull;
      keep = fm & act;
      if (keep) {
        Fn[v] = keep;
        deg = bdeg(p, v);
        red[kRMfS + x] += (unsigned long long)deg;
      }
    }
    push_frontier(p, ap, x, nx, keep != 0, v, deg);
  }
}

// Bridge step: before a side expands, test the edges between the two sides'
// current frontiers directly. A query still live after the commit at levels
// (i, j) has distance > i + j; an edge (a, b) with a in side s's frontier
// and b in side t's, both carrying the query's bit, gives a path of length
// i + 1 + j, so the distance is exactly i + j + 1 (the commit's "odd" meet
// rule). The pairs come from the two frontier lists: tiles of 64 x 64 list
// entries, one tile per wave at a time, with side t's entry broadcast
// across the lanes; pairs with a common live bit are queued in the wave's
// LDS and tested together, each lane searching the shorter id-sorted list of
// up to four pairs at once. At R-MAT s22 about two thirds of the queries end
// at distance 3, and the level-1 hubs that the next expansion would scan
// (millions of edges) belong almost only to them: the queries that survive
// a bridge carry about 5% of those edges. A pair left untested only leaves
// its queries to the expansion, which finds the same distance, so each wave
// tests at most kBridgeMaxTests pairs and the step's time stays bounded.
constexpr uint32_t kBridgeQueue = 256;             // pairs queued per wave (16 bytes each, 4 KB)
constexpr uint64_t kBridgeMaxPairs = 1ull << 26;   // list-entry pairs a bridge step may enumerate
constexpr uint32_t kBridgeMaxTests = 1024;         // pairs a wave tests (16 rounds of 64 searches)
constexpr uint64_t kBridgeMinEdges = 1ull << 20;   // live frontier edges of the cheaper side (default)
static_assert(kBridgeQueue * 16 * kWavesPerBlock <= 2 * kWavesPerBlock * kBStage * sizeof(int32_t),
              "bridge queues fit the appends' stage");

// Test the queued pairs (wave-collective): OR the masks of adjacent pairs
// into meets. Each lane takes the pairs lane, lane + 64, ... of the queue
// together, and searches the id-sorted list of each pair's lower-degree end
// for the other end. A round samples kSearchWays entries of every open range
// at once, so it costs one memory round trip for all of the lane's pairs and
// shrinks each range about kSearchWays times: a list of n entries takes
// about log8(n) rounds, where a binary search took log2(n) rounds of one
// dependent load each. The level-1 frontiers of R-MAT queries are mostly
// hubs, and the lower degree of a tested pair is typically a few hundred.
constexpr int kSearchWays = 8;

template <typename OffT>
__device__ __forceinline__ void bridge_flush(const BatchParams<OffT>& p, const int32_t* qa, const int32_t* qb,
                                             const unsigned long long* qm, uint32_t cnt,
                                             unsigned long long& meets) {
  // pairs a lane searches at once: the whole queue for 32-bit offsets
  constexpr int P = sizeof(OffT) == 4 ? (int)(kBridgeQueue / kWave) : 2;
  __builtin_amdgcn_wave_barrier();
  const uint32_t lane = (uint32_t)lane_id();
  for (uint32_t base = 0; base < cnt; base += (uint32_t)P * kWave) {
    OffT lo[P], hi[P];
    int32_t key[P];
#pragma unroll
    for (int r = 0; r < P; ++r) {
      const uint32_t k = base + (uint32_t)r * kWave + lane;
      const bool on = k < cnt;
      const int32_t a = on ? qa[k] : 0, b = on ? qb[k] : 0;
      const OffT a0 = p.off[a], a1 = p.off[a + 1], b0 = p.off[b], b1 = p.off[b + 1];
      const bool sa = a1 - a0 <= b1 - b0;
      lo[r] = sa ? a0 : b0;
      hi[r] = on ? (sa ? a1 : b1) : lo[r];
      key[r] = sa ? b : a;
    }
    uint32_t found = 0;
    // Every round at least divides each open range by kSearchWays and a
    // range of at most kSearchWays entries is read whole, so the loop ends.
    while (true) {
      bool open = false;
#pragma unroll
      for (int r = 0; r < P; ++r) open |= lo[r] < hi[r];
      if (!ballot(open)) break;
      int32_t v[P][kSearchWays];
      OffT step[P];
#pragma unroll
      for (int r = 0; r < P; ++r) {
        const OffT n = hi[r] - lo[r];
        step[r] = n <= kSearchWays ? (OffT)1 : (n + kSearchWays - 1) / kSearchWays;
        if (n > 0) {
#pragma unroll
          for (int i = 0; i < kSearchWays; ++i) {
            const OffT pos = lo[r] + (OffT)i * step[r];
            v[r][i] = p.col[pos < hi[r] ? pos : hi[r] - 1];
          }
        }
      }
#pragma unroll
      for (int r = 0; r < P; ++r) {
        if (lo[r] >= hi[r]) continue;
        int c = 0;
        bool hit = false;
#pragma unroll
        for (int i = 0; i < kSearchWays; ++i) {
          if (lo[r] + (OffT)i * step[r] < hi[r]) {
            hit |= v[r][i] == key[r];
            c += v[r][i] < key[r] ? 1 : 0;
          }
        }
        if (hit) found |= 1u << r;
        if (hit || step[r] == 1 || c == 0) {
          hi[r] = lo[r];
        } else {
          // samples 0 .. c-1 are below the key and sample c (if any) above
          // it: the key can only lie strictly between samples c-1 and c
          const OffT nlo = lo[r] + (OffT)(c - 1) * step[r];
          const OffT nhi = nlo + step[r];
          lo[r] = nlo + 1;
          hi[r] = nhi < hi[r] ? nhi : hi[r];
        }
      }
    }
#pragma unroll
    for (int r = 0; r < P; ++r)
      if (found & (1u << r)) meets |= qm[base + (uint32_t)r * kWave + lane];
  }
  __builtin_amdgcn_wave_barrier();
}

template <typename OffT>
__device__ __forceinline__ void bridge_pass(const BatchParams<OffT>& p, unsigned long long act, int gwave,
                                            int nwaves, unsigned char* lds, unsigned long long& meets) {
  const BatchCtl* ctl = p.ctl;
  const int lane = lane_id(), w = (int)(threadIdx.x >> 6);
  int32_t* qa = reinterpret_cast<int32_t*>(lds + (size_t)w * kBridgeQueue * 16);
  int32_t* qb = qa + kBridgeQueue;
  unsigned long long* qm = reinterpret_cast<unsigned long long*>(qb + kBridgeQueue);
  const int c0 = ctl->cur[0], c1 = ctl->cur[1];
  const uint32_t V = (uint32_t)p.V;
  const uint32_t n0 = ctl->fl_n[0][c0] < V ? ctl->fl_n[0][c0] : V;
  const uint32_t n1 = ctl->fl_n[1][c1] < V ? ctl->fl_n[1][c1] : V;
  uint32_t h0 = (uint32_t)(ctl->hub[0][c0] >> kHubEdgeBits), h1 = (uint32_t)(ctl->hub[1][c1] >> kHubEdgeBits);
  h0 = h0 < p.hcap ? h0 : p.hcap;
  h1 = h1 < p.hcap ? h1 : p.hcap;
  const uint32_t nS = n0 + h0, nT = n1 + h1;
  const uint32_t ct = (nT + kWave - 1) / kWave;
  const uint64_t tiles = (uint64_t)((nS + kWave - 1) / kWave) * ct;
  uint32_t cnt = 0, tested = 0;
  for (uint64_t tile = (uint64_t)gwave; tile < tiles && tested < kBridgeMaxTests; tile += (uint64_t)nwaves) {
    // tiles < 2^32: a bridge step runs only when the lists' product is at
    // most kBridgeMaxPairs
    const uint32_t ti = (uint32_t)tile / ct;
    const uint32_t i = ti * kWave + (uint32_t)lane;
    const uint32_t j = ((uint32_t)tile - ti * ct) * kWave + (uint32_t)lane;
    int32_t a = -1, b = -1;
    unsigned long long fa = 0, fb = 0;
    if (i < nS) {
      a = i < n0 ? p.fl[0][c0][i] : p.hl[0][c0][i - n0];
      if ((uint32_t)a < V)
        fa = p.F[0][c0][a] & act;
      else
        atomicOr(&p.ctl->err, 2u);
    }
    if (j < nT) {
      b = j < n1 ? p.fl[1][c1][j] : p.hl[1][c1][j - n1];
      if ((uint32_t)b < V)
        fb = p.F[1][c1][b] & act;
      else
        atomicOr(&p.ctl->err, 2u);
    }
    uint64_t cols = ballot(fb != 0);
    if (!ballot(fa != 0)) cols = 0;
    while (cols) {
      const int k = __builtin_ctzll(cols);
      cols &= cols - 1;
      const unsigned long long m = fa & (unsigned long long)bcast64((int64_t)fb, k);
      const int32_t bk = bcast(b, k);
      const bool hit = m != 0 && a != bk;
      const uint64_t hm = ballot(hit);
      if (!hm) continue;
      if (cnt + kWave > kBridgeQueue) {
        bridge_flush(p, qa, qb, qm, cnt, meets);
        tested += cnt;
        cnt = 0;
        if (tested >= kBridgeMaxTests) break;
      }
      if (hit) {
        const uint32_t pos = cnt + lane_rank(hm);
        qa[pos] = a;
        qb[pos] = bk;
        qm[pos] = m;
      }
      cnt += (uint32_t)__popcll(hm);
    }
  }
  if (cnt) bridge_flush(p, qa, qb, qm, cnt, meets);
}

// Whether a bridge step runs before the pending step decision (lane 0 of the
// finalising wave): not under the plain schedule, which is the reference's
// order of work, and once per pair of levels.
template <typename OffT>
__device__ __forceinline__ bool bridge_due(const BatchParams<OffT>& p, const BatchCtl* ctl);

// Features of the pending step for the learned scheduler (12 values in
// about [0, 1]): per side the live frontier's size and degree sum relative
// to the graph's (on a log scale), its mean degree relative to the graph's,
// the side's level and the fraction of (vertex, query) pairs it has visited;
// the graph's mean excess degree sum d^2 / sum d (the expected degree of a
// vertex reached over an edge, which sets how fast frontiers grow, on a log
// scale); the live-query fraction.
__device__ __forceinline__ void batch_features(BatchCtl* ctl, const uint64_t* mf, int64_t V, int64_t E) {
  auto lg = [](double x) { return log2(x + 1.0); };
  float* f = ctl->feat;
  const int built[2] = {(int)__popcll(ctl->lact[0]), (int)__popcll(ctl->lact[1])};
  const int live[2] = {(int)__popcll(ctl->lact[0] & ctl->act), (int)__popcll(ctl->lact[1] & ctl->act)};
  double n[2];
  for (int x = 0; x < 2; ++x) {
    const int c = ctl->cur[x];
    n[x] = (double)ctl->fl_n[x][c] + (double)(ctl->hub[x][c] >> kHubEdgeBits);
    if (built[x] > 0) n[x] *= (double)live[x] / (double)built[x];
  }
  const double pairs = (double)(ctl->live0 > 0 ? ctl->live0 : 1) * (double)V;
  const double lv = lg((double)V), le = lg((double)E), avg = (double)E / (double)(V > 0 ? V : 1);
  f[0] = (float)(lg(n[0]) / lv);
  f[1] = (float)(lg((double)mf[0]) / le);
  f[2] = (float)(lg(n[1]) / lv);
  f[3] = (float)(lg((double)mf[1]) / le);
  f[4] = ctl->f_branch;
  f[5] = (float)((double)ctl->vis_pairs[0] / pairs);
  f[6] = (float)(ctl->lvl[0] < 64 ? ctl->lvl[0] : 64) / 64.f;
  f[7] = (float)(ctl->lvl[1] < 64 ? ctl->lvl[1] : 64) / 64.f;
  f[8] = (float)(lg(n[0] > 0 ? (double)mf[0] / n[0] / avg : 0.0) / 16.0);
  f[9] = (float)(lg(n[1] > 0 ? (double)mf[1] / n[1] / avg : 0.0) / 16.0);
  f[10] = (float)__popcll(ctl->act) / 64.f;
  f[11] = (float)((double)ctl->vis_pairs[1] / pairs);
}

// The smallest power of two in [4, 64] that is at least the mean degree
// ceil(mf / n): vw < ceil(mf / n) exactly when vw * n < mf, so no division.
__device__ __forceinline__ int bdegree_width(uint64_t n, uint64_t mf) {
  if (n == 0) return 4;
  int vw = 4;
  while (vw < 64 && (uint64_t)vw * n < mf) vw <<= 1;
  return vw;
}

__device__ __forceinline__ void batch_apply_action(BatchDecision& d, int a) {
  const int k = a % 3, kind = a / 3;
  d.sides = k == 0 ? kSideS : k == 1 ? kSideT : kSideBoth;
  d.strat = kind == 2 ? kBottomUp : kTopDown;
  d.action = a;
}

// Heuristic: expand both sides while the two frontiers together have at most
// this many edges, about the work an extra step's fixed cost would buy
// (default; FASTCON_BOTH_MAX_EDGES overrides it).
constexpr uint64_t kBothMaxEdges = 1u << 20;
// After a bridge, both sides expand together only while neither live
// frontier has more than this many times the other's edges. Measured: s22
// (1,024 pairs) 2.06 ms at 4, 2.08 at 16 and 2.08 with no bound; s24
// (16,384 pairs, 30% with an endpoint outside the giant component) 89.6,
// 107.6 and 134.8 ms. Never expanding both: s22 2.31 ms, s24 96.5 ms. A
// bound of 2 gave s24 88.0 ms but s22 2.15 ms; 8 gave 96.1 and 2.04 ms.
constexpr uint64_t kBothMaxRatio = 4;

// floor(x / d) for x < 2^53 and 1 <= d <= 64 without a 64-bit integer
// division (a long instruction sequence on this chip): the double quotient is
// exact to within 2^-53 relative, so its truncation is off by at most one,
// which one comparison corrects.
__device__ __forceinline__ uint64_t div_small(uint64_t x, uint32_t d) {
  uint64_t q = (uint64_t)((double)x / (double)d);
  if (q * d > x)
    --q;
  else if ((q + 1) * d <= x)
    ++q;
  return q;
}

// Degree sum of side x's current frontier restricted to the live queries,
// estimated from the fraction of the queries the lists were built for that
// are still live (exact right after a filter).
__device__ __forceinline__ uint64_t live_mf(const BatchCtl* ctl, int x) {
  const uint64_t mf = ctl->fl_mf[x][ctl->cur[x]];
  const int built = __popcll(ctl->lact[x]);
  const int live = __popcll(ctl->lact[x] & ctl->act);
  if (built == 0 || live == built) return mf;
  return div_small(mf * (uint64_t)live, (uint32_t)built);  // mf < 2^40 and live <= 64: no overflow
}

template <typename OffT>
__device__ __forceinline__ bool bridge_due(const BatchParams<OffT>& p, const BatchCtl* ctl) {
  if (ctl->policy == 0 || !p.sorted_lists || !ctl->act) return false;
  const int32_t key = ((ctl->lvl[0] + 1) << 16) | (ctl->lvl[1] + 1);
  if (ctl->bridge_lv == key) return false;
  uint64_t n[2];
  for (int x = 0; x < 2; ++x) {
    const int c = ctl->cur[x];
    n[x] = (uint64_t)ctl->fl_n[x][c] + (ctl->hub[x][c] >> kHubEdgeBits);
  }
  if (!n[0] || !n[1] || n[0] * n[1] > kBridgeMaxPairs) return false;
  const uint64_t m0 = live_mf(ctl, 0), m1 = live_mf(ctl, 1);
  return (m0 < m1 ? m0 : m1) >= ctl->bridge_min_e;
}

// Heuristic (or plain) action for the pending step; also fills the feature
// vector f when the learned policy or a trace needs it (double-precision
// logarithms on one lane: not worth computing otherwise). One lane of the
// finalising wave.
__device__ __forceinline__ int batch_heuristic(BatchCtl* ctl, int64_t V, int64_t E, float* f, bool feats) {
  uint64_t mf[2];
  for (int x = 0; x < 2; ++x) mf[x] = live_mf(ctl, x);
  if (feats) {
    batch_features(ctl, mf, V, E);
    for (int k = 0; k < 12; ++k) f[k] = ctl->feat[k];
  }
  if (ctl->policy == 0) return 2;  // plain: both sides top-down
  if (mf[0] + mf[1] <= ctl->both_max_e) return 2;
  // Right after a bridge step at these levels (i, j), no live query has
  // distance i + j + 1 unless the bridge left pairs untested, so the next
  // answers need both sides to advance: expanding one side first would only
  // add a commit and a filter. Both sides, unless one frontier is large
  // enough for bottom-up, or more than kBothMaxRatio times the other: then
  // the small side alone is cheap, and the queries still live are often
  // ones whose small side is about to run out (an endpoint outside the
  // giant component), which ends them without the large expansion.
  const uint64_t lo = mf[0] < mf[1] ? mf[0] : mf[1], hi = mf[0] < mf[1] ? mf[1] : mf[0];
  if (ctl->bridge_lv == (((ctl->lvl[0] + 1) << 16) | (ctl->lvl[1] + 1)) && hi * 4 <= (uint64_t)E &&
      hi <= kBothMaxRatio * lo)
    return 2;
  const int side = mf[0] <= mf[1] ? 0 : 1;
  // A bit-parallel bottom-up vertex stops only once it has found every live
  // query's bit: worth it only when the top-down frontier would touch a large
  // part of the graph anyway.
  return side + (mf[side] * 4 > (uint64_t)E ? 6 : 0);
}

// Turn action a into the step decision and the next phase, filtering first
// the lists the step will read when they carry resolved queries and are
// large enough for that to pay.
__device__ __forceinline__ void batch_apply(BatchCtl* ctl, int a, bool allow_filter) {
  BatchDecision d;
  batch_apply_action(d, a);
  uint64_t n = 0, m = 0;
  for (int x = 0; x < 2; ++x) {
    if (!(d.sides & (1 << x))) continue;
    const int c = ctl->cur[x];
    n += ctl->fl_n[x][c];
    m += ctl->fl_mf[x][c] - (ctl->hub[x][c] & kHubEdgeMask);
  }
  d.vw = bdegree_width(n, m);
  if (ctl->vw_fixed > 0) d.vw = ctl->vw_fixed;
  if (a / 3 == 1) d.vw = 64;  // the wide top-down actions
  ctl->dec = d;
  int filt = 0;
  if (allow_filter) {
    for (int x = 0; x < 2; ++x) {
      if (!(d.sides & (1 << x))) continue;
      const int c = ctl->cur[x];
      // expected work saved: the share dead / built of the list's queries
      // resolved since it was built (integer products: the lane-0 decision
      // sits on every decision launch's critical path, and double-precision
      // divisions there are slow)
      const uint64_t built = (uint64_t)__popcll(ctl->lact[x]);
      const uint64_t dead = (uint64_t)__popcll(ctl->lact[x] & ~ctl->act);
      if (!dead || !built) continue;
      const uint64_t size = (uint64_t)ctl->fl_n[x][c] + (ctl->hub[x][c] >> kHubEdgeBits);
      if (size * dead >= ctl->filt_min_v * built || ctl->fl_mf[x][c] * dead >= ctl->filt_min_e * built)
        filt |= 1 << x;
    }
  }
  if (filt) {
    ctl->filt = filt;
    ctl->phase = kBFilter;
    return;
  }
  ctl->phase = d.strat == kBottomUp ? kBBottomUp : kBExpand;
}

// Return codes of batch_finalize: what the finalising wave does next.
enum : int { kFinDone = 0, kFinDecide = 1, kFinDecideFilter = 2 };

// After a step: clean up when the wave is done, else ask for a decision.
__device__ __forceinline__ int batch_next(BatchCtl* ctl, bool allow_filter) {
  if (!ctl->act) {
    ctl->phase = kBCleanup;
    return kFinDone;
  }
  return allow_filter ? kFinDecideFilter : kFinDecide;
}

// Counters a launch updated with atomics, as the finalisation needs them:
// INIT's lists at parity 0 (and the live mask), the other phases' next
// lists. Read from the finalising wave's copy of the control block, which
// was loaded after every block had voted.
struct FinPre {
  uint32_t ln[2];
  unsigned long long lh[2], lm[2], act0;
  uint32_t next_wave;  // cleanup: the wave claimed from the shared counter
};

// Distances a finalisation resolves: base + 1 for the queries in odd, base +
// 2 for those in even, -1 (not connected) for those in dead. The whole
// finalising wave writes them, one lane per query.
struct FinDist {
  unsigned long long odd, even, dead;
  int base;
};

__device__ __forceinline__ FinPre fin_prefetch(const BatchCtl* ctl, int phase, const int* nxt) {
  FinPre f{{0u, 0u}, {0ull, 0ull}, {0ull, 0ull}, 0ull, 0u};
  if (phase == kBInit) {
    f.act0 = ctl->act;
#pragma unroll
    for (int x = 0; x < 2; ++x) {
      f.ln[x] = ctl->fl_n[x][0];
      f.lh[x] = ctl->hub[x][0];
      f.lm[x] = ctl->fl_mf[x][0];
    }
  } else if (phase == kBCommit || phase == kBBottomUp || phase == kBFilter) {
#pragma unroll
    for (int x = 0; x < 2; ++x) {
      f.ln[x] = ctl->fl_n[x][nxt[x]];
      f.lh[x] = ctl->hub[x][nxt[x]];
    }
  }
  return f;
}

// ctl: the finalising wave's copy of the control block's state (LDS).
template <typename OffT>
__device__ __forceinline__ int batch_finalize(const BatchParams<OffT>& p, BatchCtl* ctl, int phase,
                                              const BatchDecision& d, const RedTot& r, const FinPre& f,
                                              FinDist& fd) {
  const uint32_t ln[2] = {f.ln[0], f.ln[1]};
  const unsigned long long lh[2] = {f.lh[0], f.lh[1]}, lm[2] = {f.lm[0], f.lm[1]}, act0 = f.act0;
  ctl->active_vertices += (int64_t)r.v[kRActV];
  ctl->active_edges += (int64_t)r.v[kRActE];
  ctl->touched_total += (int64_t)r.v[kRTouched];
  ctl->vis_pairs[0] += (int64_t)r.v[kRVisS];
  ctl->vis_pairs[1] += (int64_t)r.v[kRVisT];
  switch (phase) {
    case kBInit: {
      const unsigned long long act = act0;
      ctl->act = act;
      ctl->lvl[0] = ctl->lvl[1] = 0;
      ctl->live0 = __popcll(act);
      ctl->vis_pairs[0] = ctl->vis_pairs[1] = ctl->live0;
      ctl->bridge_lv = 0;
      for (int x = 0; x < 2; ++x) {
        ctl->fl_n[x][0] = ln[x];
        ctl->hub[x][0] = lh[x];
        ctl->fl_mf[x][0] = lm[x];
        ctl->lact[x] = act;
      }
      return batch_next(ctl, false);
    }
    case kBExpand:
      ctl->phase = kBCommit;
      return kFinDone;
    case kBCommit:
    case kBBottomUp: {
      const unsigned long long act = ctl->act;
      const unsigned long long odd = r.v[kROdd] & act;
      const unsigned long long even = r.v[kREven] & act & ~odd;
      unsigned long long alive = act;
      if (d.sides & kSideS) alive &= r.v[kRAnyS];
      if (d.sides & kSideT) alive &= r.v[kRAnyT];
      const unsigned long long dead = act & ~alive & ~odd & ~even;
      fd = FinDist{odd, even, dead, ctl->lvl[0] + ctl->lvl[1]};
      ctl->act = act & ~(odd | even | dead);
      ctl->steps++;
      ctl->levels_total++;
      for (int x = 0; x < 2; ++x) {
        if (!(d.sides & (1 << x))) continue;
        const int c = ctl->cur[x];
        ctl->lvl[x]++;
        ctl->lact[x] = act;
        // The consumed parity becomes the next "next": its counters restart.
        // After a bottom-up step its lists are still needed to clear the
        // masks, so kBClearF restarts them.
        if (phase == kBCommit) {
          ctl->fl_n[x][c] = 0;
          ctl->hub[x][c] = 0;
        }
        ctl->fl_mf[x][c] = 0;
        ctl->cur[x] = c ^ 1;
        ctl->fl_n[x][c ^ 1] = ln[x];
        ctl->hub[x][c ^ 1] = lh[x];
        ctl->fl_mf[x][c ^ 1] = r.v[kRMfS + x];
      }
      if (phase == kBBottomUp) {
        ctl->bu_sides = d.sides;
        ctl->phase = kBClearF;  // the consumed masks are cleared first
        return kFinDone;
      }
      return batch_next(ctl, true);
    }
    case kBBridge: {
      const unsigned long long act = ctl->act;
      const unsigned long long hit = r.v[kROdd] & act;
      fd = FinDist{hit, 0ull, 0ull, ctl->lvl[0] + ctl->lvl[1]};
      ctl->act = act & ~hit;
      ctl->bridges++;
      ctl->bridged += __popcll(hit);
      return batch_next(ctl, true);
    }
    case kBClearF: {
      for (int x = 0; x < 2; ++x) {
        if (!(ctl->bu_sides & (1 << x))) continue;
        ctl->fl_n[x][ctl->cur[x] ^ 1] = 0;
        ctl->hub[x][ctl->cur[x] ^ 1] = 0;
      }
      return batch_next(ctl, true);
    }
    case kBFilter: {
      const unsigned long long act = ctl->act;
      for (int x = 0; x < 2; ++x) {
        if (!(ctl->filt & (1 << x))) continue;
        const int c = ctl->cur[x];
        ctl->fl_n[x][c] = 0;
        ctl->hub[x][c] = 0;
        ctl->fl_mf[x][c] = 0;
        ctl->cur[x] = c ^ 1;
        ctl->fl_n[x][c ^ 1] = ln[x];
        ctl->hub[x][c ^ 1] = lh[x];
        ctl->fl_mf[x][c ^ 1] = r.v[kRMfS + x];
        ctl->lact[x] = act;
      }
      ctl->filt = 0;
      ctl->filters++;
      // The decision is taken again on the exact live frontiers and may pick
      // a side that was not filtered: that side may be filtered in turn. A
      // filtered side holds no resolved query, so at most two filters run.
      return batch_next(ctl, true);
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
      ctl->act = 0;
      ctl->waves_done++;
      // the next wave: claimed from the counter the concurrent machines
      // share (so that a machine whose waves ran fast takes more of them),
      // or this machine's next one
      ctl->wave = p.wave_ticket ? (int32_t)f.next_wave : ctl->wave + 1;
      ctl->t_wave_end = (int64_t)wall_clock64();
      ctl->phase = ctl->wave < ctl->nwaves && ctl->wave < ctl->stop_wave ? kBInit : kBIdle;
      return kFinDone;
    }
    default:
      ctl->phase = kBIdle;
  }
  return kFinDone;
}

// Seed a wave (block 0, waves 0 and 1 = sides s and t, one lane per query).
// Records, masks and the dirty bitmap of the roots are zero here, so plain
// stores suffice; repeated roots within a side are merged in LDS first. The
// counters are written with agent-scope stores for the finalising block.
template <typename OffT>
__device__ __forceinline__ void init_wave(const BatchParams<OffT>& p, int32_t* s_root) {
  BatchCtl* ctl = p.ctl;
  const int q = lane_id(), x = (int)(threadIdx.x >> 6);
  const int qi = ctl->wave * 64 + q;
  bool live = false;
  int32_t s = 0, t = 0;
  if (qi < ctl->nq) {
    s = p.qs[qi];
    t = p.qt[qi];
    live = s != t && (uint32_t)s < (uint32_t)p.V && (uint32_t)t < (uint32_t)p.V;
    if (x == 0 && s == t) p.dist[qi] = 0;
    if (live && p.perm) {
      s = p.perm[s];
      t = p.perm[t];
    }
  }
  const uint64_t lm = ballot(live);
  const int32_t u = live ? (x ? t : s) : -1;
  s_root[x * 64 + q] = u;
  __builtin_amdgcn_wave_barrier();
  unsigned long long mask = 0;
  bool owner = live;
  if (live) {
    for (int j = 0; j < 64; ++j) {
      if (s_root[x * 64 + j] == u) {
        mask |= 1ull << j;
        if (j < q) owner = false;
      }
    }
  }
  int64_t deg = 0;
  if (owner) {
    rec_of(p, x, u)[0] = mask;
    p.F[x][0][u] = mask;
    atomicOr(&p.db[u >> 5], 1u << (u & 31));
    deg = bdeg(p, u);
  }
  const bool hub = owner && deg >= p.hub_threshold;
  const uint64_t nm = ballot(owner && !hub), hm = ballot(hub);
  if (owner && !hub) p.fl[x][0][lane_rank(nm)] = u;
  const unsigned long long hd = hub ? (unsigned long long)deg : 0ull;
  const unsigned long long incl = wave_incl_scan_u64(hd);
  if (hub) {
    const uint32_t slot = lane_rank(hm);
    if (slot < p.hcap) {
      p.hl[x][0][slot] = u;
      p.hb[x][0][slot] = (int64_t)(incl - hd);
    }
  }
  const unsigned long long mf = wave_sum_u64(owner ? (unsigned long long)deg : 0ull);
  const unsigned long long hsum = wave_sum_u64(hd);
  if (q == 0) {
    st_agent(&ctl->fl_n[x][0], (uint32_t)__popcll(nm));
    const uint32_t nh = (uint32_t)__popcll(hm);
    if (nh > p.hcap) atomicOr(&ctl->err, 1u);
    st_agent(&ctl->hub[x][0], ((unsigned long long)nh << kHubEdgeBits) | hsum);
    st_agent(&ctl->fl_mf[x][0], mf);
    if (x == 0) st_agent(&ctl->act, (unsigned long long)lm);
  }
}

template <typename OffT>
__global__ __launch_bounds__(kBlock, 4) void batch_kernel(BatchParams<OffT> p) {
  BatchCtl* ctl = p.ctl;
  const int phase = ctl->phase;
  if (phase == kBIdle || phase == kBPause) return;
  const uint32_t nb = gridDim.x;
  const int gwave = (int)(blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6));
  const int nwaves = (int)(nb * kWavesPerBlock);
  const int64_t gtid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t nthr = (int64_t)nb * kBlock;
  __shared__ int32_t s_stage[2][kWavesPerBlock * kBStage];
  __shared__ int32_t s_hub[2][2][kWavesPerBlock * kHubStage];
  __shared__ BShared sh;
  const BatchDecision d = ctl->dec;
  const unsigned long long act = ctl->act;
  const int w = (int)(threadIdx.x >> 6);
  const int nxt[2] = {ctl->cur[0] ^ 1, ctl->cur[1] ^ 1};
  Appends ap{{{&s_stage[0][w * kBStage], 0u}, {&s_stage[1][w * kBStage], 0u}},
             {{&s_hub[0][0][w * kHubStage], &s_hub[0][1][w * kHubStage], 0u},
              {&s_hub[1][0][w * kHubStage], &s_hub[1][1][w * kHubStage], 0u}}};
  unsigned long long red[kRCount];
#pragma unroll
  for (int k = 0; k < kRCount; ++k) red[k] = 0;

  switch (phase) {
    case kBInit: {
      if (blockIdx.x == 0 && threadIdx.x < 128) init_wave(p, &s_stage[0][0]);
      break;
    }
    case kBExpand: {
      const int chunk = ctl->tile > 0 ? ctl->tile : 256;
#pragma unroll
      for (int x = 0; x < 2; ++x)
        if (d.sides & (1 << x)) expand_side(p, x, d.vw, chunk, act, gwave, nwaves, red[kRActV], red[kRActE]);
      block_reduce_slots(red, p.red, sh);
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
      commit_pass(p, ap, act, nxt, gwave, nwaves, red);
      appends_block_flush(p, ap, nxt, sh);
      block_reduce_slots(red, p.red, sh);
      break;
    }
    case kBBottomUp: {
      if (d.sides == kSideS)
        bottom_up<OffT, 1>(p, ap, act, gwave, nwaves, red);
      else if (d.sides == kSideT)
        bottom_up<OffT, 2>(p, ap, act, gwave, nwaves, red);
      else
        bottom_up<OffT, 3>(p, ap, act, gwave, nwaves, red);
      appends_block_flush(p, ap, nxt, sh);
      block_reduce_slots(red, p.red, sh);
      break;
    }
    case kBFilter: {
      if (ctl->filt & 1) filter_side<OffT, 0>(p, ap, act, gwave, nwaves, red);
      if (ctl->filt & 2) filter_side<OffT, 1>(p, ap, act, gwave, nwaves, red);
      appends_block_flush(p, ap, nxt, sh);
      block_reduce_slots(red, p.red, sh);
      break;
    }
    case kBBridge: {
      bridge_pass(p, act, gwave, nwaves, reinterpret_cast<unsigned char*>(&s_stage[0][0]), red[kROdd]);
      block_reduce_slots(red, p.red, sh);
      break;
    }
    case kBClearF: {
      for (int x = 0; x < 2; ++x) {
        if (!(ctl->bu_sides & (1 << x))) continue;
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
      }
      break;
    }
    case kBCleanup: {
      // The only frontier masks still set are those of each side's current
      // lists (an expansion clears the masks it reads, a commit its hub
      // masks, a filter and kBClearF the lists they consume): clear them
      // through the lists. Every vertex with a non-zero record is dirty:
      // clear the records of the dirty vertices in address order.
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
      const int lane = lane_id();
      const int64_t per = owned_groups(((int64_t)p.V + 31) / 32, gwave, nwaves);
      for (int64_t jb = 0; jb < per; jb += kWave) {
        const GroupDeal gd{jb, nwaves, gwave};
        const int64_t wi = gd(lane);
        const uint32_t word = jb + lane < per ? p.db[wi] : 0u;
        if (word) p.db[wi] = 0u;
        uint64_t nz = ballot(word != 0);
        while (nz) {
          const int k0 = __builtin_ctzll(nz);
          nz &= nz - 1;
          int k1 = -1;
          if (nz) {
            k1 = __builtin_ctzll(nz);
            nz &= nz - 1;
          }
          const int mine = lane < 32 ? k0 : k1;
          const uint32_t wv = (uint32_t)__shfl((int)word, mine < 0 ? 0 : mine);
          const int64_t u = gd(mine < 0 ? 0 : mine) * 32 + (lane & 31);
          if (mine >= 0 && ((wv >> (lane & 31)) & 1u) && u < p.V) {
            *reinterpret_cast<ulonglong2*>(rec_of(p, 0, u)) = make_ulonglong2(0ull, 0ull);
            *reinterpret_cast<ulonglong2*>(rec_of(p, 1, u)) = make_ulonglong2(0ull, 0ull);
          }
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
    // Two-level ticket: the blocks of group b % kRedSlots count on their own
    // word, and the last block of each group counts on the launch ticket, so
    // no single address takes an atomic from every block.
    const uint32_t k = blockIdx.x % kRedSlots;
    const uint32_t gsize = (nb - k + kRedSlots - 1) / kRedSlots;
    unsigned long long* sub = p.red + (size_t)(kRedSlots + k) * kRedWords;
    const unsigned long long t1 = __hip_atomic_fetch_add(sub, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    bool last = false;
    if (t1 == gsize - 1) {
      __hip_atomic_store(sub, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const uint32_t groups = nb < (uint32_t)kRedSlots ? nb : (uint32_t)kRedSlots;
      const uint32_t t2 = __hip_atomic_fetch_add(&ctl->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      last = t2 == groups - 1;
    }
    sh.last = last;
  }
  __syncthreads();
  if (!sh.last || threadIdx.x >= kWave) return;
  // Wave 0 of the last block finalises: slot sums, the state update (lane
  // 0), and, when a step must be chosen, the decision (the learned policy
  // evaluates its network across the wave). The network's weights are
  // copied from global memory straight into LDS (the appends' stage is free
  // now) first, so that the copy overlaps the finalisation.
  constexpr int kPer = (kMlpParams + kWave - 1) / kWave;
  static_assert(kPer * kWave * sizeof(float) <= sizeof(s_stage), "network does not fit the LDS stage");
  const bool mlp_on = ctl->policy == 2 && p.mlp != nullptr;
  float* s_mlp = reinterpret_cast<float*>(&s_stage[0][0]);
  if (mlp_on) {
    typedef __attribute__((address_space(1))) void* gptr;
    typedef __attribute__((address_space(3))) void* lptr;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      // lane l writes word k * kWave + l of s_mlp; lanes past the end re-read
      // the last weight into the unused tail of the stage
      const int i = k * kWave + lane_id();
      const float* src = p.mlp + (i < kMlpParams ? i : kMlpParams - 1);
      __builtin_amdgcn_global_load_lds((gptr)src, (lptr)(s_mlp + k * kWave), 4, 0, 0);
    }
  }
  // The state update and the decision run on an LDS copy of the control
  // block's state (everything before the phase log), loaded with one
  // agent-scope load per lane and written back with one store per lane: one
  // lane's chain of read-modify-writes on global fields had cost a round
  // trip each (loads and stores share one wait counter), 4-11 us per launch.
  // Every other block has voted, so no one else writes these words now.
  // The state words and the launch's reduction slots are loaded together,
  // in one round trip.
  __shared__ __align__(16) unsigned long long s_ctl[kCtlHotWords];
  // A cleanup claims the machine's next wave now, so that the counter's
  // round trip overlaps the state loads.
  uint32_t next_wave = 0;
  if (phase == kBCleanup && p.wave_ticket && lane_id() == 0) next_wave = atomicAdd(p.wave_ticket, 1u);
  RedTot tot;
  {
    const int lane = lane_id();
    const unsigned long long* gw = reinterpret_cast<const unsigned long long*>(ctl);
    unsigned long long c0 = 0, c1 = 0;
    if (lane < (int)kCtlHotWords) c0 = ld_agent(gw + lane);
    if (lane + kWave < (int)kCtlHotWords) c1 = ld_agent(gw + lane + kWave);
    unsigned long long* row = p.red + (size_t)(lane < kRedSlots ? lane : 0) * kRedWords;
    unsigned long long v[kRCount];
#pragma unroll
    for (int k = 0; k < kRCount; ++k) v[k] = lane < kRedSlots ? ld_agent(row + k) : 0ull;
    if (lane < (int)kCtlHotWords) s_ctl[lane] = c0;
    if (lane + kWave < (int)kCtlHotWords) s_ctl[lane + kWave] = c1;
#pragma unroll
    for (int k = 0; k < kRCount; ++k) {
      if (v[k]) row[k] = 0ull;
      tot.v[k] = k < kRMfS ? (unsigned long long)wave_reduce_full_nz_u64<kWaveOr>(v[k])
                           : (unsigned long long)wave_reduce_full_nz_u64<kWaveSum>(v[k]);
    }
  }
  __builtin_amdgcn_wave_barrier();
  BatchCtl* lc = reinterpret_cast<BatchCtl*>(s_ctl);
  int need = kFinDone;
  {
    FinPre pre = fin_prefetch(lc, phase, nxt);
    pre.next_wave = next_wave;
    const int wave0 = lc->wave * 64;
    FinDist fd{0ull, 0ull, 0ull, 0};
    if (threadIdx.x == 0) {
      const int64_t ph = lc->phases;
      if (ph < kPhaseLog) {
        ctl->plog[ph] = (int8_t)(phase * 16 + (phase == kBExpand || phase == kBBottomUp ? d.sides : 0));
        ctl->ptime[ph] = (int64_t)wall_clock64();
        ctl->pwork[ph] = (int64_t)(phase == kBCommit ? tot.v[kRTouched] : tot.v[kRActE]);
      }
      lc->phases = ph + 1;
      if (phase == kBDecide) {
        need = lc->pause_filter ? kFinDecideFilter : kFinDecide;
      } else {
        need = batch_finalize(p, lc, phase, d, tot, pre, fd);
        if (need != kFinDone && bridge_due(p, lc)) {
          lc->bridge_lv = ((lc->lvl[0] + 1) << 16) | (lc->lvl[1] + 1);
          lc->phase = kBBridge;
          need = kFinDone;
        }
      }
    }
    // the resolved queries' distances, one lane per query
    const unsigned long long odd = (unsigned long long)bcast64((int64_t)fd.odd, 0);
    const unsigned long long even = (unsigned long long)bcast64((int64_t)fd.even, 0);
    const unsigned long long dead = (unsigned long long)bcast64((int64_t)fd.dead, 0);
    const int base = bcast(fd.base, 0);
    const int q = lane_id();
    const unsigned long long b = 1ull << q;
    if ((odd | even | dead) & b) p.dist[wave0 + q] = (odd & b) ? base + 1 : (even & b) ? base + 2 : -1;
  }
  need = __shfl(need, 0);
  if (need != kFinDone) {
    float f[12] = {};
    int a = 0;
    if (threadIdx.x == 0) a = batch_heuristic(lc, p.V, p.E, f, mlp_on || lc->trace);
    a = __shfl(a, 0);
    const bool forced = phase == kBDecide && lc->force_action >= 0;
    const bool pause = phase != kBDecide && lc->trace &&
                       lc->decisions % (lc->trace_every > 0 ? lc->trace_every : 1) == 0;
    if (!forced && !pause && mlp_on) {
      __builtin_amdgcn_s_waitcnt(0);  // the weights are in LDS
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int k = 0; k < 12; ++k) f[k] = __shfl(f[k], 0);
      a = mlp_pick_wave_t(s_mlp, f, a, 1e-3f * (float)lc->learn_margin_milli);
    }
    if (threadIdx.x == 0) {
      if (pause) {
        lc->heur_action = a;
        lc->pause_filter = need == kFinDecideFilter;
        lc->phase = kBPause;
      } else {
        if (forced) a = lc->force_action;
        batch_apply(lc, a, need == kFinDecideFilter);
        lc->decisions++;
      }
      if (phase == kBDecide) lc->t_decide = (int64_t)wall_clock64();
    }
  }
  // no copy into LDS may still be in flight when the wave ends
  if (mlp_on) __builtin_amdgcn_s_waitcnt(0);
  if (threadIdx.x == 0) lc->ticket = 0u;
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
  {
    unsigned long long* gw = reinterpret_cast<unsigned long long*>(ctl);
    for (int k = lane_id(); k < (int)kCtlHotWords; k += kWave) gw[k] = s_ctl[k];
  }
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
  uint8_t* tm = nullptr;
  unsigned long long* red = nullptr;
  int32_t* fl[2][2] = {};
  int32_t* hl[2][2] = {};
  int64_t* hb[2][2] = {};
  uint32_t* db = nullptr;
  BatchCtl* ctl = nullptr;
  int32_t* qs = nullptr;
  int32_t* qt = nullptr;
  int32_t* dist = nullptr;
  uint32_t* wticket = nullptr;  // next unclaimed wave (lane 0's is shared by all lanes of a run)
  int32_t qcap = 0;
  int blocks = 0;
  // Query list, results and wave counter of the current run: lane 0's for
  // every lane. A captured graph holds the pointers it was captured with.
  const int32_t* run_qs = nullptr;
  const int32_t* run_qt = nullptr;
  int32_t* run_dist = nullptr;
  uint32_t* run_ticket = nullptr;
  const void* cap_key[4] = {};
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int graph_force = -2;
  int tuned_vw = -1;       // autotuner result (0 = degree-chosen)
  int tuned_tile = -1;
  double est_launches = 12.0;  // working launches per wave (running estimate)
  int32_t* h_dist = nullptr;  // pinned
  int32_t* h_phase = nullptr; // pinned
  char* snap = nullptr;        // tracing: copy of the workspace
  // Concurrent lanes: lane 0 runs on the graph's stream; each further lane
  // (a wave machine of its own, on its own stream) hangs off `next`.
  hipStream_t stream = nullptr;
  bool own_stream = false;
  hipEvent_t ev_end = nullptr;
  BatchWorkspace* next = nullptr;
};

void batch_free(BatchWorkspace* b) {
  if (!b) return;
  batch_free(b->next);
  if (b->own_stream && b->stream) {
    (void)hipStreamSynchronize(b->stream);
    (void)hipStreamDestroy(b->stream);
  }
  if (b->ev_end) (void)hipEventDestroy(b->ev_end);
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  (void)hipFree(b->mem);
  (void)hipFree(b->qs);
  (void)hipFree(b->qt);
  (void)hipFree(b->dist);
  if (b->h_dist) (void)hipHostFree(b->h_dist);
  if (b->h_phase) (void)hipHostFree(b->h_phase);
  (void)hipFree(b->snap);
  delete b;
}

static fc_status batch_alloc_lane(fc_graph* g, BatchWorkspace** out, bool own_stream) {
  BatchWorkspace* b = new BatchWorkspace();
  const size_t V = (size_t)g->V;
  const size_t hcap = g->hcap;
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  size_t o = 0;
  const size_t o_rec = o;
  o += al(32 * V);
  const size_t o_F = o;
  o += 4 * al(8 * V);
  const size_t o_tm = o;
  o += al((V + 31) / 32 + 64);
  const size_t o_red = o;
  o += al(sizeof(unsigned long long) * 2 * kRedSlots * kRedWords);  // slots, then group tickets
  const size_t o_fl = o;
  o += 4 * al(4 * V);
  const size_t o_hl = o;
  o += 4 * al(4 * hcap);
  const size_t o_hb = o;
  o += 4 * al(8 * hcap);
  const size_t o_db = o;
  o += al(4 * ((V + 63) / 32));
  const size_t o_ctl = o;
  o += al(sizeof(BatchCtl));
  const size_t o_wt = o;
  o += al(sizeof(uint32_t));
  b->bytes = o;
  hipError_t e = hipMalloc((void**)&b->mem, b->bytes);
  if (e != hipSuccess) {
    delete b;
    return hip_fail(e, "batch workspace allocation");
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
  b->tm = (uint8_t*)(b->mem + o_tm);
  b->red = (unsigned long long*)(b->mem + o_red);
  b->db = (uint32_t*)(b->mem + o_db);
  b->ctl = (BatchCtl*)(b->mem + o_ctl);
  b->wticket = (uint32_t*)(b->mem + o_wt);
  int occ = g->off64 ? batch_occupancy<int64_t>() : batch_occupancy<int32_t>();
  if (occ < 1) {
    batch_free(b);
    return fail(FC_ERR_NOGPU, "batch kernel has no occupancy on this device");
  }
  // With three concurrent lanes, 2 blocks per CU per lane measured best at
  // s22 (1: 6.19 ms, 2: 5.58, 3: 5.79, 4: 6.26 for 1,024 queries).
  int per_cu = std::min(occ, 2);
  const char* bpc = std::getenv("FASTCON_BATCH_BLOCKS_PER_CU");
  if (bpc && std::atoi(bpc) > 0) per_cu = std::min(occ, std::atoi(bpc));
  b->blocks = per_cu * g->cus;
  if (hipHostMalloc((void**)&b->h_phase, 4 * sizeof(int32_t), hipHostMallocDefault) != hipSuccess) {
    batch_free(b);
    return fail(FC_ERR_NOMEM, "pinned allocation failed");
  }
  if (own_stream) {
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
      b->stream = nullptr;
      batch_free(b);
      return fail(FC_ERR_HIP, "batch lane stream creation failed");
    }
    b->own_stream = true;
  } else {
    b->stream = g->stream;
  }
  // Cleared on the lane's own stream and finished before the first use: the
  // lanes' streams do not wait for the legacy null stream, and a plain
  // hipMemset there may still be running when the lane writes its control
  // block (the last bytes of the allocation).
  e = hipMemsetAsync(b->mem, 0, b->bytes, b->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  if (e != hipSuccess) {
    batch_free(b);
    return hip_fail(e, "batch workspace clear");
  }
  if (hipEventCreate(&b->ev_end) != hipSuccess) {
    b->ev_end = nullptr;
    batch_free(b);
    return fail(FC_ERR_HIP, "batch lane event creation failed");
  }
  *out = b;
  return FC_OK;
}

static fc_status batch_alloc(fc_graph* g) {
  if (g->batch) return FC_OK;
  return batch_alloc_lane(g, &g->batch, false);
}

// The k-th lane (0 = g->batch), allocated on first use.
static fc_status batch_lane(fc_graph* g, int k, BatchWorkspace** out) {
  fc_status st = batch_alloc(g);
  if (st != FC_OK) return st;
  BatchWorkspace* b = g->batch;
  for (int i = 0; i < k; ++i) {
    if (!b->next) {
      st = batch_alloc_lane(g, &b->next, true);
      if (st != FC_OK) return st;
    }
    b = b->next;
  }
  *out = b;
  return FC_OK;
}

static fc_status batch_queries(fc_graph* g, BatchWorkspace* b, int32_t n) {
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
static BatchParams<OffT> batch_params(fc_graph* g, BatchWorkspace* b) {
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
  p.tm = b->tm;
  p.red = b->red;
  p.db = b->db;
  p.ctl = b->ctl;
  p.qs = b->run_qs;
  p.qt = b->run_qt;
  p.dist = b->run_dist;
  p.wave_ticket = b->run_ticket;
  p.perm = g->relabeled ? g->d_perm : nullptr;
  p.v_active = (int32_t)g->v_active;
  p.sorted_lists = g->lists_sorted ? 1 : 0;
  p.mlp = g->d_mlp + kMlpParams;  // the copy in mlp_t_index order
  return p;
}

static hipError_t batch_launch_raw(fc_graph* g, BatchWorkspace* b) {
  if (g->off64) return launch_batch_phase<int64_t>(batch_params<int64_t>(g, b), b->blocks, b->stream);
  return launch_batch_phase<int32_t>(batch_params<int32_t>(g, b), b->blocks, b->stream);
}

// Point a lane at the run's query list, results and wave counter.
static void batch_bind(BatchWorkspace* b, const int32_t* qs, const int32_t* qt, int32_t* dist, uint32_t* ticket) {
  b->run_qs = qs;
  b->run_qt = qt;
  b->run_dist = dist;
  b->run_ticket = ticket;
}

static fc_status batch_graph(fc_graph* g, BatchWorkspace* b) {
  const void* key[4] = {b->run_qs, b->run_qt, b->run_dist, b->run_ticket};
  if (b->gexec && std::memcmp(key, b->cap_key, sizeof(key)) == 0) return FC_OK;
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
  std::memcpy(b->cap_key, key, sizeof(key));
  FC_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  hipError_t le = hipSuccess;
  for (int i = 0; i < kGraphLaunches && le == hipSuccess; ++i) le = batch_launch_raw(g, b);
  hipGraph_t gr = nullptr;
  hipError_t ce = hipStreamEndCapture(b->stream, &gr);
  if (le != hipSuccess) return hip_fail(le, "batch kernel capture");
  if (ce != hipSuccess) return hip_fail(ce, "hipStreamEndCapture");
  b->graph = gr;
  FC_HIP(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
  return FC_OK;
}

// Prepare the control block for a batch of n queries (already on the device).
fc_status batch_begin(fc_graph* g, BatchWorkspace* b, int32_t n, int policy, int vw, int tile, int trace, int every,
                      float margin, int32_t first_wave) {
  BatchCtl init;
  std::memset(&init, 0, sizeof(init));
  init.phase = kBInit;
  init.wave = first_wave;
  init.nwaves = (n + 63) / 64;
  init.nq = n;
  init.policy = policy;
  init.vw_fixed = vw;
  init.tile = tile;
  init.force_action = -1;
  init.stop_wave = init.nwaves;
  init.trace = trace;
  init.trace_every = every;
  init.learn_margin_milli = (int32_t)std::lround((margin < 0.f ? kDefaultLearnMargin : margin) * 1000.f);
  init.f_branch = g->f_branch;
  // Thresholds are capped at 2^56 (beyond any graph's counts), so that the
  // decision's products with query counts (at most 64) cannot overflow.
  auto env_u64 = [](const char* name, uint64_t dflt) {
    const char* v = std::getenv(name);
    const uint64_t x = v && std::strtoull(v, nullptr, 10) > 0 ? (uint64_t)std::strtoull(v, nullptr, 10) : dflt;
    return std::min<uint64_t>(x, 1ull << 56);
  };
  init.filt_min_v = env_u64("FASTCON_FILTER_MIN_VERTICES", kFilterMinVertices);
  init.filt_min_e = env_u64("FASTCON_FILTER_MIN_EDGES", kFilterMinEdges);
  init.both_max_e = env_u64("FASTCON_BOTH_MAX_EDGES", kBothMaxEdges);
  init.bridge_min_e = env_u64("FASTCON_BRIDGE_MIN_EDGES", kBridgeMinEdges);
  FC_HIP(hipMemcpyAsync(b->ctl, &init, sizeof(init), hipMemcpyHostToDevice, b->stream));
  // the reduction slots are cleared by every finalisation; clear them here
  // too in case an earlier batch stopped on an error
  FC_HIP(hipMemsetAsync(b->red, 0, sizeof(unsigned long long) * 2 * kRedSlots * kRedWords, b->stream));
  return FC_OK;
}

// Run the phase machines of `nl` lanes until every one is idle, issuing each
// lane's launches on its own stream (use_graph: replay captured launches).
// The elapsed time runs from an event on lane 0 that the other lanes wait
// for to the point where every lane has finished.
struct LaneRun {
  BatchWorkspace* b;
  int64_t nw;        // waves this lane is expected to take
  int64_t nw_total;  // waves of the batch
  int64_t expected;
  int64_t launched;
  int round;
  bool done;
};

static fc_status batch_drive(fc_graph* g, LaneRun* lr, int nl, bool use_graph, float* ms, int64_t* launches_out) {
  BatchWorkspace* b0 = lr[0].b;
  FC_HIP(hipEventRecord(g->ev0, b0->stream));
  for (int k = 1; k < nl; ++k) FC_HIP(hipStreamWaitEvent(lr[k].b->stream, g->ev0, 0));
  for (int k = 0; k < nl; ++k) {
    // 10% above the expected working launches: a lane that runs out waits
    // for the host's next poll, which costs more than a few no-op launches
    lr[k].expected = (int64_t)std::ceil((double)lr[k].nw * lr[k].b->est_launches * 1.1) + 2;
    lr[k].launched = 0;
    lr[k].round = 0;
    lr[k].done = lr[k].nw == 0;
  }
  for (;;) {
    bool any = false;
    for (int k = 0; k < nl; ++k) {
      LaneRun& L = lr[k];
      if (L.done) continue;
      any = true;
      const int64_t want = std::max<int64_t>(L.expected - L.launched, 1);
      if (!use_graph || want < kGraphLaunches) {
        for (int64_t i = 0; i < want; ++i) {
          hipError_t e = batch_launch_raw(g, L.b);
          if (e != hipSuccess) return hip_fail(e, "batch kernel launch");
        }
        L.launched += want;
      } else {
        const int64_t ng = (want + kGraphLaunches - 1) / kGraphLaunches;
        for (int64_t i = 0; i < ng; ++i) FC_HIP(hipGraphLaunch(L.b->gexec, L.b->stream));
        L.launched += ng * kGraphLaunches;
      }
      FC_HIP(hipEventRecord(L.b->ev_end, L.b->stream));
      FC_HIP(hipMemcpyAsync(L.b->h_phase, L.b->ctl, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, L.b->stream));
    }
    if (!any) break;
    for (int k = 0; k < nl; ++k) {
      LaneRun& L = lr[k];
      if (L.done) continue;
      FC_HIP(hipStreamSynchronize(L.b->stream));
      if (L.b->h_phase[0] == kBIdle) {
        L.done = true;
        continue;
      }
      if (L.launched > L.nw_total * (4 * g->V + 16) + 64)
        return fail(FC_ERR_CAPACITY, "batch phase machine did not finish after %lld launches", (long long)L.launched);
      // Waves are claimed dynamically, so a lane's remaining work is not
      // known: budget one more wave, then more each round.
      L.expected = L.launched + (int64_t)std::ceil((L.b->est_launches + 4.0) * (1.0 + L.round)) + 2;
      L.round++;
    }
  }
  for (int k = 1; k < nl; ++k) FC_HIP(hipStreamWaitEvent(b0->stream, lr[k].b->ev_end, 0));
  FC_HIP(hipEventRecord(g->ev1, b0->stream));
  FC_HIP(hipEventSynchronize(g->ev1));
  FC_HIP(hipEventElapsedTime(ms, g->ev0, g->ev1));
  int64_t total = 0;
  for (int k = 0; k < nl; ++k) total += lr[k].launched;
  if (launches_out) *launches_out = total;
  return FC_OK;
}

// Answer n queries with `lanes` concurrent wave machines. Every lane reads
// lane 0's query list and writes lane 0's results; lane k starts with wave
// k and each lane, when it finishes a wave, claims the next unclaimed one
// from a counter in lane 0's workspace, so lanes that meet cheap waves take
// more of them. Distances land in lane 0's pinned h_dist[0..n).
static fc_status batch_run(fc_graph* g, const int32_t* s, const int32_t* t, int32_t n, int policy, int vw, int tile,
                           bool use_graph, float* ms, BatchCtl* ctl_out, int64_t* launches = nullptr,
                           float margin = -1.f, int lanes = 1) {
  const int64_t nw = (n + 63) / 64;
  const int nl = (int)std::max<int64_t>(1, std::min<int64_t>(lanes, nw));
  LaneRun lr[kMaxLanes];
  fc_status st;
  BatchWorkspace* b0 = nullptr;
  if ((st = batch_lane(g, 0, &b0)) != FC_OK) return st;
  if ((st = batch_queries(g, b0, n)) != FC_OK) return st;
  FC_HIP(hipMemcpyAsync(b0->qs, s, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, b0->stream));
  FC_HIP(hipMemcpyAsync(b0->qt, t, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, b0->stream));
  const uint32_t first_unclaimed = (uint32_t)nl;
  if (nl > 1)
    FC_HIP(hipMemcpyAsync(b0->wticket, &first_unclaimed, sizeof(uint32_t), hipMemcpyHostToDevice, b0->stream));
  for (int k = 0; k < nl; ++k) {
    BatchWorkspace* b = nullptr;
    if ((st = batch_lane(g, k, &b)) != FC_OK) return st;
    batch_bind(b, b0->qs, b0->qt, b0->dist, nl > 1 ? b0->wticket : nullptr);
    if (use_graph && (st = batch_graph(g, b)) != FC_OK) return st;
    if ((st = batch_begin(g, b, n, policy, vw, tile, 0, 1, margin, k)) != FC_OK) return st;
    lr[k].b = b;
    lr[k].nw = (nw + nl - 1 - k) / nl;
    lr[k].nw_total = nw;
  }
  // the copies above read pageable host memory and the lanes read lane 0's
  // arrays: every stream is idle before the first launch
  for (int k = 0; k < nl; ++k) FC_HIP(hipStreamSynchronize(lr[k].b->stream));
  if ((st = batch_drive(g, lr, nl, use_graph, ms, launches)) != FC_OK) return st;
  std::vector<BatchCtl> cs(nl);
  FC_HIP(hipMemcpyAsync(b0->h_dist, b0->dist, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost, b0->stream));
  for (int k = 0; k < nl; ++k)
    FC_HIP(hipMemcpyAsync(&cs[k], lr[k].b->ctl, sizeof(BatchCtl), hipMemcpyDeviceToHost, lr[k].b->stream));
  for (int k = 0; k < nl; ++k) FC_HIP(hipStreamSynchronize(lr[k].b->stream));
  uint32_t err = 0;
  for (int k = 0; k < nl; ++k) err |= cs[k].err;
  if (err) return fail(FC_ERR_CAPACITY, "batch kernel: list guard tripped (err=%u)", err);
  int64_t done = 0;
  for (int k = 0; k < nl; ++k) done += cs[k].waves_done;
  if (done != nw)
    return fail(FC_ERR_HIP, "batch: %lld of %lld waves finished (lane 0: phase %d, wave %d of %d, %lld working launches)",
                (long long)done, (long long)nw, cs[0].phase, cs[0].wave, cs[0].nwaves, (long long)cs[0].phases);
  // Running estimate of the working launches (finalised phases) per wave.
  // It follows the phases, not the launches issued: those are rounded up to
  // whole captured graphs, and an estimate fed by them would only grow.
  for (int k = 0; k < nl; ++k)
    if (cs[k].waves_done > 0)
      lr[k].b->est_launches = 0.5 * lr[k].b->est_launches + 0.5 * ((double)cs[k].phases / (double)cs[k].waves_done);
  if (ctl_out) {
    *ctl_out = cs[0];
    for (int k = 1; k < nl; ++k) {
      ctl_out->levels_total += cs[k].levels_total;
      ctl_out->active_vertices += cs[k].active_vertices;
      ctl_out->active_edges += cs[k].active_edges;
      ctl_out->touched_total += cs[k].touched_total;
      ctl_out->filters += cs[k].filters;
      ctl_out->phases += cs[k].phases;
      ctl_out->bridges += cs[k].bridges;
      ctl_out->bridged += cs[k].bridged;
    }
  }
  if (std::getenv("FASTCON_BATCH_LOG")) {
    // phase sequence of lane 0: I=init E=expand(s/t/b sides) C=commit U=bottom-up X=clear F=filter Z=cleanup
    // D=traced decision B=bridge
    const BatchCtl& c = cs[0];
    const char* nm = "?IECUXZFPDB?????";
    const int64_t np = std::min<int64_t>(c.phases, kPhaseLog);
    for (int64_t i = 0; i < np; ++i) {
      const int ph = (uint8_t)c.plog[i] >> 4, sd = c.plog[i] & 15;
      const double us = i ? (double)(c.ptime[i] - c.ptime[i - 1]) / 100.0 : 0.0;  // 100 MHz clock
      std::fprintf(stderr, "%c%s:%.0f", nm[ph & 15], sd == 1 ? "s" : sd == 2 ? "t" : sd == 3 ? "b" : "", us);
      if (c.pwork[i]) std::fprintf(stderr, "/%lldk", (long long)(c.pwork[i] / 1000));
      std::fputc(ph == kBCleanup ? '\n' : ' ', stderr);
    }
  }
  return FC_OK;
}

// Autotuner: time the first waves of this batch under each (width, tile)
// candidate (the median of three runs) and keep the fastest for this graph.
static fc_status batch_autotune(fc_graph* g, const int32_t* s, const int32_t* t, int32_t n, int policy, int lanes) {
  BatchWorkspace* b = g->batch;
  // Eight waves, run as the batch itself runs (concurrent lanes, replayed
  // graphs): two waves on one lane with direct launches had the tile choice
  // flip from run to run.
  const int32_t m = std::min<int32_t>(n, 512);
  const int vws[] = {0, 8, 32, 64};
  const int tiles[] = {256, 1024};
  float best = 1e30f;
  int bvw = 0, btile = 256;
  for (int vw : vws)
    for (int tile : tiles) {
      // median of three timings, so that one noisy run does not pick the pair
      float ts[3];
      for (float& x : ts) {
        BatchCtl c;
        fc_status st = batch_run(g, s, t, m, policy, vw, tile, true, &x, &c, nullptr, -1.f, lanes);
        if (st != FC_OK) return st;
      }
      std::sort(ts, ts + 3);
      const float ms = ts[1];
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

// Tracing: run direct launches until the phase machine pauses or goes idle.
static fc_status drive_to_stop(fc_graph* g, int64_t cap) {
  BatchWorkspace* b = g->batch;
  for (int64_t launched = 0;; launched += 8) {
    for (int i = 0; i < 8; ++i) {
      hipError_t e = batch_launch_raw(g, b);
      if (e != hipSuccess) return hip_fail(e, "batch kernel launch");
    }
    FC_HIP(hipMemcpyAsync(b->h_phase, b->ctl, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, b->stream));
    FC_HIP(hipStreamSynchronize(b->stream));
    if (b->h_phase[0] == kBIdle || b->h_phase[0] == kBPause) return FC_OK;
    if (launched > cap) return fail(FC_ERR_CAPACITY, "batch phase machine did not stop after %lld launches", (long long)launched);
  }
}

static fc_status put_ctl(fc_graph* g, const BatchCtl& c) {
  FC_HIP(hipMemcpyAsync(g->batch->ctl, &c, sizeof(BatchCtl), hipMemcpyHostToDevice, g->stream));
  FC_HIP(hipStreamSynchronize(g->stream));
  return FC_OK;
}

static fc_status get_ctl(fc_graph* g, BatchCtl* c) {
  FC_HIP(hipMemcpyAsync(c, g->batch->ctl, sizeof(BatchCtl), hipMemcpyDeviceToHost, g->stream));
  FC_HIP(hipStreamSynchronize(g->stream));
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
  if (o.tile != 0 && (o.tile < 64 || o.tile > 65536 || (o.tile % 64) != 0))
    return fc::fail(FC_ERR_INVALID, "fc_stcon_batch: tile must be 0 or a multiple of 64 in [64, 65536]");
  if (o.lanes < 0 || o.lanes > fcd::kMaxLanes)
    return fc::fail(FC_ERR_INVALID, "fc_stcon_batch: lanes must be 0 (default) or 1..%d", fcd::kMaxLanes);
  const int lanes = o.lanes ? o.lanes : fcd::kDefaultLanes;
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (n == 0) return FC_OK;
  fc_status st = fcd::batch_alloc(g);
  if (st != FC_OK) return st;
  g->levels_valid = g->parents_valid = false;
  fcd::BatchWorkspace* b = g->batch;
  int vw = o.vw, tile = o.tile;
  if (vw == 0 && tile == 0) {
    if (b->tuned_vw < 0) {
      st = fcd::batch_autotune(g, s, t, n, o.policy, lanes);
      if (st != FC_OK) return st;
    }
    vw = b->tuned_vw;
    tile = b->tuned_tile;
  }
  if (tile == 0) tile = 256;
  float ms = 0;
  BatchCtl c;
  int64_t launches = 0;
  st = fcd::batch_run(g, s, t, n, o.policy, vw, tile, o.use_graph != 0, &ms, &c, &launches, o.learn_margin, lanes);
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
    stats->filters = (int32_t)c.filters;
    stats->phases = (int32_t)c.phases;
    stats->launches = launches;
    stats->bridges = c.bridges;
    stats->bridged = (int32_t)c.bridged;
  }
  return FC_OK;
}

// Trace of the heuristic schedule: at every step decision the workspace is
// copied, each action is taken from that state and the rest of the wave run
// (heuristic schedule), and the device clock gives the time from the
// decision to the end of the wave; then the copy is restored and the
// heuristic's own action taken.
extern "C" fc_status fc_sched_trace(fc_graph* g, const int32_t* s, const int32_t* t, int32_t n, int32_t policy,
                                    int32_t every,
                                    fc_trace_row* rows, int32_t cap, int32_t* n_rows, int32_t* dist_out) {
  fc::clear_error();
  if (!g || !n_rows || (n > 0 && (!s || !t)) || (cap > 0 && !rows))
    return fc::fail(FC_ERR_INVALID, "fc_sched_trace: null argument");
  if (n < 0 || cap < 0 || every < 1) return fc::fail(FC_ERR_INVALID, "fc_sched_trace: bad count or stride");
  if (policy != FC_POLICY_HEURISTIC && policy != FC_POLICY_LEARNED)
    return fc::fail(FC_ERR_INVALID, "fc_sched_trace: policy must be heuristic or learned");
  if (policy == FC_POLICY_LEARNED && !g->mlp_set)
    return fc::fail(FC_ERR_INVALID, "fc_sched_trace: learned policy requested but no scheduler parameters are set");
  *n_rows = 0;
  for (int32_t i = 0; i < n; ++i)
    if (s[i] < 0 || s[i] >= g->V || t[i] < 0 || t[i] >= g->V)
      return fc::fail(FC_ERR_INVALID, "fc_sched_trace: query %d vertex out of range", i);
  if (n == 0) return FC_OK;
  fc_status st = fcd::batch_alloc(g);
  if (st != FC_OK) return st;
  g->levels_valid = g->parents_valid = false;
  fcd::BatchWorkspace* b = g->batch;
  if ((st = fcd::batch_queries(g, b, n)) != FC_OK) return st;
  if (!b->snap && hipMalloc((void**)&b->snap, b->bytes) != hipSuccess) {
    b->snap = nullptr;
    return fc::fail(FC_ERR_NOMEM, "fc_sched_trace: no room for the workspace copy (%zu bytes)", b->bytes);
  }
  int khz = 0;
  FC_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, g->device));
  if (khz <= 0) return fc::fail(FC_ERR_HIP, "fc_sched_trace: device clock rate unknown");
  FC_HIP(hipMemcpyAsync(b->qs, s, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, g->stream));
  FC_HIP(hipMemcpyAsync(b->qt, t, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, g->stream));
  fcd::batch_bind(b, b->qs, b->qt, b->dist, nullptr);
  const int vw = b->tuned_vw >= 0 ? b->tuned_vw : 0;
  const int tile = b->tuned_tile > 0 ? b->tuned_tile : 256;
  if ((st = fcd::batch_begin(g, b, n, policy, vw, tile, cap > 0 ? 1 : 0, every, 0.f, 0)) != FC_OK) return st;
  const int64_t cap_launches = ((int64_t)n / 64 + 1) * (4 * g->V + 16) + 64;
  std::vector<BatchCtl> cv(2);
  BatchCtl& c = cv[0];
  BatchCtl& r = cv[1];
  for (;;) {
    if ((st = fcd::drive_to_stop(g, cap_launches)) != FC_OK) return st;
    if ((st = fcd::get_ctl(g, &c)) != FC_OK) return st;
    if (c.err) return fc::fail(FC_ERR_CAPACITY, "fc_sched_trace: list guard tripped (err=%u)", c.err);
    if (c.phase == fcd::kBIdle) break;
    fc_trace_row row;
    std::memset(&row, 0, sizeof(row));
    for (int k = 0; k < FC_SCHED_FEATURES; ++k) row.features[k] = c.feat[k];
    row.heuristic = c.heur_action;
    row.wave = c.wave;
    FC_HIP(hipMemcpyAsync(b->snap, b->mem, b->bytes, hipMemcpyDeviceToDevice, g->stream));
    for (int a = 0; a < FC_SCHED_ACTIONS; ++a) {
      if (a > 0) FC_HIP(hipMemcpyAsync(b->mem, b->snap, b->bytes, hipMemcpyDeviceToDevice, g->stream));
      r = c;
      r.phase = fcd::kBDecide;
      r.force_action = a;
      r.trace = 0;
      r.stop_wave = c.wave + 1;
      if ((st = fcd::put_ctl(g, r)) != FC_OK) return st;
      if ((st = fcd::drive_to_stop(g, cap_launches)) != FC_OK) return st;
      if ((st = fcd::get_ctl(g, &r)) != FC_OK) return st;
      if (r.err) return fc::fail(FC_ERR_CAPACITY, "fc_sched_trace: list guard tripped (err=%u)", r.err);
      row.cost_us[a] = (float)((double)(r.t_wave_end - r.t_decide) * 1000.0 / (double)khz);
    }
    FC_HIP(hipMemcpyAsync(b->mem, b->snap, b->bytes, hipMemcpyDeviceToDevice, g->stream));
    if (*n_rows < cap) rows[(*n_rows)++] = row;
    r = c;
    r.phase = fcd::kBDecide;
    r.force_action = -1;
    r.trace = *n_rows < cap ? 1 : 0;  // full: run the rest untraced
    if ((st = fcd::put_ctl(g, r)) != FC_OK) return st;
  }
  if (dist_out) {
    FC_HIP(hipMemcpyAsync(dist_out, b->dist, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost, g->stream));
    FC_HIP(hipStreamSynchronize(g->stream));
  }
  return FC_OK;
}
