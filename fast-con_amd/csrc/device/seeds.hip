// Helper-seed s–t connectivity: the multi-source single-frontier visit of
// BFS_4K_MULTI (BFS_4K_MULTI/src/device/BFS_WorkEfficient.cu:9-168,
// .../WorkEfficientKernel/BFS_WE_Kernels1.cu:25-48) and its colour-contact
// closure (.../WorkEfficientKernel/BFS_WE_AdjMatrixAnalysis.cu:1-35).
//
// Colour 0 grows from s, colour 1 from t and colours 2..K-1 from random
// helper seeds, all in one frontier. A vertex is claimed once, with one
// atomicCAS of (colour << kSeedShift) | level, as in the ONE_ATOMIC s–t
// visit; the reference writes the colour without an atomic. Every probe that
// meets another colour sets that pair's bits in a K x K contact matrix.
// After each level one block closes the matrix from colour 0. s and t are
// connected as soon as the closure reaches colour 1. They are not connected
// when the frontier empties first: every seed's component is then covered
// by the balls of the seeds inside it, and two balls of one component touch
// along some edge. The answer is a verdict; the path through helper balls is
// no shortest path, so no distance comes out.
//
// Launches follow the engine's pattern: the host enqueues (level, check)
// pairs in batches, and a pair issued after the search ended is a no-op.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <unordered_set>
#include <vector>

#include "common.hpp"
#include "engine.hpp"

using fc::fail;

namespace fcd {

constexpr int kSeedShift = 22;                        // state = colour << 22 | level
constexpr int32_t kSeedLevelMask = (1 << kSeedShift) - 1;
constexpr int kMaxSeeds = 512;                        // colours (fits bits 22..30)
constexpr int kSeedWords = kMaxSeeds / 32;            // contact-row words

struct SeedsCtl {
  int32_t level;     // frontier = queue[level & 1], its vertices at this level
  int32_t done;      // 0 running, 1 connected, 2 frontier empty, 3 level limit
  uint32_t qn[2];    // queue sizes by parity
  uint32_t err;
  int32_t pad;
  int64_t claimed;   // vertices claimed after the seeds
  uint32_t reach[kSeedWords];  // colours joined to colour 0 so far (only grows)
};

template <typename OffT>
struct SeedsParams {
  const OffT* off;
  const int32_t* col;
  int32_t V;
  int32_t K;
  int32_t* state;
  int32_t* q[2];
  uint32_t* contact;  // [K][kSeedWords]
  SeedsCtl* ctl;
};

__global__ void seeds_fill_kernel(int32_t* __restrict__ state, int64_t V, uint32_t* __restrict__ contact,
                                  int64_t cwords) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += stride) state[v] = -1;
  for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < cwords; w += stride) contact[w] = 0u;
}

// One block: the seeds (distinct vertices) take colours 0..K-1 at level 0.
template <typename OffT>
__global__ void seeds_start_kernel(SeedsParams<OffT> p, const int32_t* __restrict__ seeds) {
  for (int i = (int)threadIdx.x; i < p.K; i += (int)blockDim.x) {
    const int32_t v = seeds[i];
    p.state[v] = i << kSeedShift;
    p.q[0][i] = v;
  }
  if (threadIdx.x == 0) {
    SeedsCtl* c = p.ctl;
    c->level = 0;
    c->done = 0;
    c->qn[0] = (uint32_t)p.K;
    c->qn[1] = 0u;
    c->err = 0u;
    c->claimed = 0;
    for (int w = 0; w < kSeedWords; ++w) c->reach[w] = w == 0 ? 1u : 0u;
  }
}

// Claimed vertices are staged per wave in LDS and published with one atomic
// on the frontier counter per full stage, instead of one per wave and edge.
constexpr uint32_t kSeedStage = 512;

struct SeedStage {
  int32_t* buf;
  uint32_t cnt;  // wave-uniform
};

__device__ __forceinline__ void seed_flush(SeedStage& st, int32_t* nq, uint32_t* nqn, uint32_t cap, uint32_t* err) {
  if (!st.cnt) return;
  __builtin_amdgcn_wave_barrier();
  uint32_t base = 0;
  if (lane_id() == 0) base = atomicAdd(nqn, st.cnt);
  base = (uint32_t)__shfl((int)base, 0);
  for (uint32_t i = (uint32_t)lane_id(); i < st.cnt; i += kWave) {
    if (base + i < cap)
      nq[base + i] = st.buf[i];
    else
      atomicOr(err, 1u);
  }
  __builtin_amdgcn_wave_barrier();
  st.cnt = 0;
}

template <typename OffT>
__device__ __forceinline__ void seeds_probe(const SeedsParams<OffT>& p, int32_t u, int32_t mine, int32_t cv,
                                            bool act, int32_t* nq, uint32_t* nqn, uint32_t& claimed,
                                            SeedStage& st) {
  bool won = false;
  if (act) {
    int32_t cur = p.state[u];
    if (cur == -1) {
      cur = atomicCAS(&p.state[u], -1, mine);
      won = cur == -1;
    }
    if (!won) {
      const int32_t cu = cur >> kSeedShift;
      if (cu != cv) {
        uint32_t* a = p.contact + (size_t)cv * kSeedWords + (cu >> 5);
        const uint32_t ab = 1u << (cu & 31);
        if (!(*a & ab)) atomicOr(a, ab);
        uint32_t* b = p.contact + (size_t)cu * kSeedWords + (cv >> 5);
        const uint32_t bb = 1u << (cv & 31);
        if (!(*b & bb)) atomicOr(b, bb);
      }
    }
  }
  const uint64_t m = ballot(won);
  if (!m) return;
  const uint32_t k = (uint32_t)__popcll(m);
  if (st.cnt + k > kSeedStage) seed_flush(st, nq, nqn, (uint32_t)p.V, &p.ctl->err);
  if (won) {
    claimed += 1;
    st.buf[st.cnt + lane_rank(m)] = u;
  }
  st.cnt += k;
}

// Top-down level: a lane per frontier vertex for its own list when it has
// at most 64 neighbours; a vertex with more is taken by the whole wave, 64
// edges at a time.
template <typename OffT>
__global__ __launch_bounds__(256) void seeds_level_kernel(SeedsParams<OffT> p) {
  SeedsCtl* ctl = p.ctl;
  if (ctl->done) return;
  const int32_t lvl = ctl->level;
  const int par = lvl & 1;
  uint32_t n = ctl->qn[par];
  if (n > (uint32_t)p.V) n = (uint32_t)p.V;
  const int32_t* q = p.q[par];
  int32_t* nq = p.q[par ^ 1];
  uint32_t* nqn = &ctl->qn[par ^ 1];
  const int lane = lane_id();
  uint32_t claimed = 0;
  __shared__ int32_t s_stage[256 / kWave][kSeedStage];
  SeedStage st{s_stage[threadIdx.x >> 6], 0u};
  const int64_t nthr = (int64_t)gridDim.x * blockDim.x;
  const int64_t nr = ((int64_t)n + kWave - 1) / kWave * kWave;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nr; i += nthr) {
    int32_t v = -1, cv = 0, mine = 0;
    OffT b = 0, e = 0;
    if (i < n) {
      v = q[i];
      if ((uint32_t)v < (uint32_t)p.V) {
        cv = p.state[v] >> kSeedShift;
        mine = (cv << kSeedShift) | (lvl + 1);
        b = p.off[v];
        e = p.off[v + 1];
      } else {
        v = -1;
        atomicOr(&ctl->err, 2u);
      }
    }
    const bool heavy = v >= 0 && e - b > (OffT)kWave;
    // light vertices: each lane its own list
    OffT j = heavy ? e : b;
    while (ballot(j < e)) {
      const bool act = j < e;
      const int32_t u = act ? p.col[j] : 0;
      seeds_probe(p, u, mine, cv, act, nq, nqn, claimed, st);
      ++j;
    }
    // heavy vertices: the whole wave per list
    uint64_t hm = ballot(heavy);
    while (hm) {
      const int h = __builtin_ctzll(hm);
      hm &= hm - 1;
      const OffT hb = (OffT)shfl64((int64_t)b, h), he = (OffT)shfl64((int64_t)e, h);
      const int32_t hmine = __shfl(mine, h), hcv = __shfl(cv, h);
      for (OffT k = hb; k < he; k += kWave) {
        const OffT x = k + (OffT)lane;
        const bool act = x < he;
        const int32_t u = act ? p.col[x] : 0;
        seeds_probe(p, u, hmine, hcv, act, nq, nqn, claimed, st);
      }
    }
  }
  seed_flush(st, nq, nqn, (uint32_t)p.V, &ctl->err);
  const uint64_t cl = wave_sum_u64(claimed);
  if (lane == 0 && cl) atomicAdd((unsigned long long*)&ctl->claimed, cl);
}

// One block after each level: close the contact relation from colour 0 and
// decide; otherwise advance the level and restart the consumed queue's
// counter, which the next level fills. Contacts only accumulate, so the
// closure resumes from the set the previous level reached and usually
// settles in a round or two.
__global__ __launch_bounds__(256) void seeds_check_kernel(uint32_t* __restrict__ contact, int K, SeedsCtl* ctl) {
  __shared__ uint32_t reach[kSeedWords];
  __shared__ int changed;
  if (ctl->done) return;
  const int tid = (int)threadIdx.x;
  if (tid < kSeedWords) reach[tid] = ctl->reach[tid];
  __syncthreads();
  for (int round = 0; round <= K; ++round) {
    if (tid == 0) changed = 0;
    __syncthreads();
    for (int c = tid; c < K; c += (int)blockDim.x) {
      if (!((reach[c >> 5] >> (c & 31)) & 1u)) continue;
      const uint32_t* row = contact + (size_t)c * kSeedWords;
      for (int w = 0; w < kSeedWords; ++w) {
        const uint32_t add = row[w] & ~reach[w];
        if (add) {
          atomicOr(&reach[w], add);
          changed = 1;
        }
      }
    }
    __syncthreads();
    if (!changed) break;
    __syncthreads();
  }
  if (tid < kSeedWords) ctl->reach[tid] = reach[tid];
  if (tid == 0) {
    const int32_t lvl = ctl->level;
    const uint32_t next = ctl->qn[(lvl & 1) ^ 1];
    if ((reach[0] >> 1) & 1u) {
      ctl->done = 1;
    } else if (next == 0) {
      ctl->done = 2;
    } else if (lvl + 1 >= kSeedLevelMask) {
      ctl->done = 3;
    } else {
      ctl->qn[lvl & 1] = 0u;
      ctl->level = lvl + 1;
    }
  }
}

}  // namespace fcd

using namespace fcd;

// Per-graph scratch of the helper-seed visit: the contact matrix, the
// control block and the seed list (allocated on first use).
struct SeedsScratch {
  uint32_t* contact = nullptr;
  SeedsCtl* ctl = nullptr;
  int32_t* seeds = nullptr;
};

template <typename OffT>
static void seeds_launch_pair(fc_graph* g, const SeedsParams<OffT>& p) {
  const int blocks = std::max(1, g->cus * 4);
  hipLaunchKernelGGL(seeds_level_kernel<OffT>, dim3(blocks), dim3(256), 0, g->stream, p);
  hipLaunchKernelGGL(seeds_check_kernel, dim3(1), dim3(256), 0, g->stream, p.contact, p.K, p.ctl);
}

template <typename OffT>
static fc_status seeds_run(fc_graph* g, const std::vector<int32_t>& seeds, int32_t* done_out, fc_stats* stats) {
  static_assert(sizeof(SeedsCtl) <= 256, "control block");
  const int K = (int)seeds.size();
  char* mem = nullptr;
  const size_t cbytes = sizeof(uint32_t) * (size_t)kMaxSeeds * kSeedWords;
  FC_HIP(hipMalloc((void**)&mem, cbytes + 256 + sizeof(int32_t) * kMaxSeeds));
  struct Free {
    char* m;
    ~Free() { (void)hipFree(m); }
  } guard{mem};
  SeedsParams<OffT> p;
  p.off = (const OffT*)g->d_off;
  p.col = g->d_col;
  p.V = (int32_t)g->V;
  p.K = K;
  p.state = g->d_state;
  p.q[0] = g->d_q[0][0];
  p.q[1] = g->d_q[0][1];
  p.contact = (uint32_t*)mem;
  p.ctl = (SeedsCtl*)(mem + cbytes);
  int32_t* d_seeds = (int32_t*)(mem + cbytes + 256);
  FC_HIP(hipMemcpyAsync(d_seeds, seeds.data(), sizeof(int32_t) * (size_t)K, hipMemcpyHostToDevice, g->stream));
  FC_HIP(hipStreamSynchronize(g->stream));
  FC_HIP(hipEventRecord(g->ev0, g->stream));
  const int fb = (int)std::min<int64_t>((g->V + 255) / 256, (int64_t)g->cus * 8);
  hipLaunchKernelGGL(seeds_fill_kernel, dim3(fb), dim3(256), 0, g->stream, g->d_state, (int64_t)g->V, p.contact,
                     (int64_t)K * kSeedWords);
  hipLaunchKernelGGL(seeds_start_kernel<OffT>, dim3(1), dim3(256), 0, g->stream, p, (const int32_t*)d_seeds);
  FC_HIP(hipGetLastError());
  SeedsCtl h;
  int64_t pairs = 0;
  for (int batch = 8;; batch = std::min(batch * 2, 256)) {
    for (int i = 0; i < batch; ++i) seeds_launch_pair(g, p);
    FC_HIP(hipGetLastError());
    pairs += batch;
    FC_HIP(hipEventRecord(g->ev1, g->stream));
    FC_HIP(hipMemcpyAsync(&h, p.ctl, sizeof(h), hipMemcpyDeviceToHost, g->stream));
    FC_HIP(hipStreamSynchronize(g->stream));
    if (h.done) break;
    if (pairs > g->V + 64)
      return fail(FC_ERR_CAPACITY, "helper-seed visit did not finish after %lld levels", (long long)pairs);
  }
  float ms = 0.f;
  FC_HIP(hipEventElapsedTime(&ms, g->ev0, g->ev1));
  if (h.err) return fail(FC_ERR_CAPACITY, "helper-seed visit: queue guard tripped (err=%u)", h.err);
  if (h.done == 3) return fail(FC_ERR_CAPACITY, "helper-seed visit: more than %d levels", kSeedLevelMask);
  *done_out = h.done;
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->time_ms = ms;
    stats->levels = h.level + 1;
    stats->visited_vertices = h.claimed + K;
  }
  return FC_OK;
}

extern "C" fc_status fc_stcon_seeds(fc_graph* g, int32_t s, int32_t t, int32_t helpers, uint64_t seed,
                                    int32_t* connected, fc_stats* stats) {
  fc::clear_error();
  if (!g || !connected) return fail(FC_ERR_INVALID, "fc_stcon_seeds: null argument");
  if (s < 0 || s >= g->V || t < 0 || t >= g->V)
    return fail(FC_ERR_INVALID, "fc_stcon_seeds: vertex out of range [0,%lld)", (long long)g->V);
  if (helpers < 0 || helpers > kMaxSeeds - 2)
    return fail(FC_ERR_INVALID, "fc_stcon_seeds: helpers must be in [0,%d]", kMaxSeeds - 2);
  g->levels_valid = g->parents_valid = false;
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (s == t) {
    *connected = 1;
    return FC_OK;
  }
  // colours 0 and 1 are s and t; helpers are distinct vertices drawn from
  // the seeded stream (fewer when the graph has fewer vertices)
  std::vector<int32_t> seeds = {g->relabeled ? g->h_perm[(size_t)s] : s, g->relabeled ? g->h_perm[(size_t)t] : t};
  std::unordered_set<int32_t> used(seeds.begin(), seeds.end());
  const uint64_t V = (uint64_t)g->V;
  for (uint64_t k = 0; (int)seeds.size() < helpers + 2 && k < 64 * (uint64_t)(helpers + 2) && used.size() < V; ++k) {
    const int32_t r = (int32_t)(((unsigned __int128)fc::stream_word(seed, 9, k) * V) >> 64);
    if (used.insert(r).second) seeds.push_back(r);
  }
  int32_t done = 0;
  fc_status st = g->off64 ? seeds_run<int64_t>(g, seeds, &done, stats) : seeds_run<int32_t>(g, seeds, &done, stats);
  if (st != FC_OK) return st;
  *connected = done == 1 ? 1 : 0;
  return FC_OK;
}
