// Graph construction on the device: seeded R-MAT edge generation and the
// canonical CSR build (symmetrise, drop self-loops and dropped pairs, sort by
// (u, v), remove duplicates). The results equal the host fc_gen_rmat and
// fc_csr_from_edges bit for bit. This replaces make_struct's Thrust pipeline
// (st-conn/make_struct.c:304-356,553-738; st-conn/cudakernel/gstruct_kernel.cu:114-145,187-235;
// st-conn/make_struct_gpufunc.cu:242-328).
//
// Each directed copy (u, v) becomes the key (u << vbits) | v. An invalid pair
// (a self-loop or a negative endpoint) gets u = V and so sorts after every
// valid key. The keys are sorted by an LSD radix sort with 8-bit digits
// written here, one pass per digit:
//   - each 4096-key tile counts its digits;
//   - one exclusive scan of the digit-major count table gives every
//     (digit, tile) its output offset;
//   - a stable scatter takes each key's rank among equal digits from 64-lane
//     match masks.
// Adjacent duplicates are then dropped, and the offsets come from degree
// counts and one more scan.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "engine.hpp"

using fc::fail;

namespace fc {
fc_status alloc_host_graph(int64_t V, int64_t E, fc_host_graph* out);
}

namespace fcd {

constexpr int kBT = 256;                   // threads per block
constexpr int kPerThread = 16;             // elements per thread and tile
constexpr int kTile = kBT * kPerThread;    // elements per tile
constexpr int kWavesBT = kBT / kWave;

// ---------------------------------------------------------------------------
// Counter-based random words, the same functions as the host generators'
// (SplitMix64 finaliser; word i of stream (seed, lane)); the R-MAT test
// compares the two generators' graphs bit for bit.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t d_mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t d_stream_word(uint64_t seed, uint64_t lane, uint64_t i) {
  return d_mix64(d_mix64(seed ^ (lane * 0xD1B54A32D192ED03ull)) + i * 0x9E3779B97F4A7C15ull);
}

// The host Scrambler's bijection of [0, 2^bits), with its constants
// computed on the host.
struct Scramble {
  uint64_t mask, m1, m2, a1, a2;
  int shift;
  __device__ __forceinline__ uint64_t operator()(uint64_t x) const {
    x = (x * m1 + a1) & mask;
    x ^= x >> shift;
    x = (x * m2 + a2) & mask;
    x ^= x >> shift;
    x = (x * m1 + a2) & mask;
    return x;
  }
};

// ---------------------------------------------------------------------------
// Scans (uint32, exclusive), three phases: tile sums, a scan of the sums
// (recursive), and a tile scan seeded with its tile's prefix.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t x) {
  const int lane = lane_id();
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)x, o);
    if (lane >= o) x += y;
  }
  return x;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t x) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) x += (uint32_t)__shfl_xor((int)x, o);
  return x;
}

// Exclusive scan of one value per thread of the block; *total gets the
// block's sum. sh holds kWavesBT words; all threads call.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t x, uint32_t* sh, uint32_t& total) {
  const int w = (int)(threadIdx.x >> 6);
  const uint32_t inc = wave_incl_scan_u32(x);
  if (lane_id() == kWave - 1) sh[w] = inc;
  __syncthreads();
  uint32_t before = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < kWavesBT; ++k) {
    const uint32_t s = sh[k];
    before += k < w ? s : 0u;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return before + inc - x;
}

__global__ __launch_bounds__(kBT) void scan_reduce_kernel(const uint32_t* __restrict__ in, int64_t n,
                                                          uint32_t* __restrict__ partial) {
  __shared__ uint32_t sh[kWavesBT];
  const int64_t base = (int64_t)blockIdx.x * kTile;
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) {
    const int64_t i = base + (int64_t)k * kBT + threadIdx.x;
    if (i < n) s += in[i];
  }
  s = wave_sum_u32(s);
  if (lane_id() == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int k = 0; k < kWavesBT; ++k) t += sh[k];
    partial[blockIdx.x] = t;
  }
}

// out may alias in: each element is read and written by the same thread.
__global__ __launch_bounds__(kBT) void scan_down_kernel(const uint32_t* in, uint32_t* out, int64_t n,
                                                        const uint32_t* __restrict__ carry_in,
                                                        uint32_t* __restrict__ total) {
  __shared__ uint32_t sh[kWavesBT];
  const int64_t base = (int64_t)blockIdx.x * kTile;
  uint32_t carry = carry_in ? carry_in[blockIdx.x] : 0u;
  for (int k = 0; k < kPerThread; ++k) {
    const int64_t i = base + (int64_t)k * kBT + threadIdx.x;
    const uint32_t x = i < n ? in[i] : 0u;
    uint32_t tot;
    const uint32_t ex = block_excl_scan(x, sh, tot);
    if (i < n) out[i] = carry + ex;
    carry += tot;
  }
  if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *total = carry;
}

static int64_t tiles_of(int64_t n) { return (n + kTile - 1) / kTile; }

// Words of scratch scan_u32 needs for n elements.
static int64_t scan_scratch(int64_t n) {
  const int64_t t = tiles_of(n);
  if (t <= 1) return 0;
  return ((t + 63) & ~(int64_t)63) + scan_scratch(t);
}

// Exclusive scan of in[0..n) into out (may alias); *total (device, nullable)
// receives the sum. tmp holds scan_scratch(n) words.
static hipError_t scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total, uint32_t* tmp,
                           hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int64_t t = tiles_of(n);
  if (t == 1) {
    hipLaunchKernelGGL(scan_down_kernel, dim3(1), dim3(kBT), 0, s, in, out, n, (const uint32_t*)nullptr, total);
    return hipGetLastError();
  }
  uint32_t* partial = tmp;
  uint32_t* rest = tmp + ((t + 63) & ~(int64_t)63);
  hipLaunchKernelGGL(scan_reduce_kernel, dim3((unsigned)t), dim3(kBT), 0, s, in, n, partial);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = scan_u32(partial, partial, t, nullptr, rest, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(scan_down_kernel, dim3((unsigned)t), dim3(kBT), 0, s, in, out, n, (const uint32_t*)partial,
                     total);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Keys
// ---------------------------------------------------------------------------
__global__ void edge_keys_kernel(const int64_t* __restrict__ edges, int64_t m, int64_t V, int vbits,
                                 uint64_t* __restrict__ keys, uint32_t* __restrict__ err) {
  const uint64_t bad = (uint64_t)V << vbits;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
    const int64_t u = edges[2 * i], v = edges[2 * i + 1];
    if (u >= V || v >= V) atomicOr(err, 1u);
    const bool ok = u >= 0 && v >= 0 && u != v && u < V && v < V;
    keys[2 * i] = ok ? ((uint64_t)u << vbits) | (uint64_t)v : bad;
    keys[2 * i + 1] = ok ? ((uint64_t)v << vbits) | (uint64_t)u : bad;
  }
}

// R-MAT edge i as the host generator draws it: one 32-bit uniform per level
// picks the quadrant against the thresholds ta <= tb <= tc, and both ids go
// through the scrambler.
__global__ void rmat_keys_kernel(int scale, int64_t M, uint64_t seed, uint64_t ta, uint64_t tb, uint64_t tc,
                                 Scramble perm, int vbits, uint64_t* __restrict__ keys) {
  const uint64_t bad = (1ull << scale) << vbits;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < M; i += stride) {
    uint64_t u = 0, v = 0, w = 0;
    for (int lvl = 0; lvl < scale; ++lvl) {
      if ((lvl & 1) == 0) w = d_stream_word(seed, 1, (uint64_t)i * 16 + (uint64_t)(lvl >> 1));
      const uint64_t r = (lvl & 1) ? (w >> 32) : (w & 0xFFFFFFFFull);
      const uint64_t bu = r >= tb, bv = (r >= ta && r < tb) || r >= tc;
      u = (u << 1) | bu;
      v = (v << 1) | bv;
    }
    u = perm(u);
    v = perm(v);
    const bool ok = u != v;
    keys[2 * i] = ok ? (u << vbits) | v : bad;
    keys[2 * i + 1] = ok ? (v << vbits) | u : bad;
  }
}

// Largest endpoint (for V = 0: max + 1, as the host builder).
__global__ void max_endpoint_kernel(const int64_t* __restrict__ edges, int64_t n2,
                                    unsigned long long* __restrict__ mx) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  long long best = -1;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += stride)
    best = edges[i] > best ? edges[i] : best;
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    const long long y = (long long)shfl64(best, lane_id() ^ o);
    best = y > best ? y : best;
  }
  if (lane_id() == 0 && best >= 0) atomicMax(mx, (unsigned long long)best + 1ull);
}

// ---------------------------------------------------------------------------
// Radix sort
// ---------------------------------------------------------------------------
// Lanes (among the valid ones) whose 8-bit digit equals this lane's.
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool valid) {
  uint64_t m = ballot(valid);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const bool bit = ((d >> b) & 1u) != 0;
    const uint64_t bb = ballot(bit);
    m &= bit ? bb : ~bb;
  }
  return m;
}

__global__ __launch_bounds__(kBT) void radix_hist_kernel(const uint64_t* __restrict__ keys, int64_t n, int shift,
                                                         uint32_t* __restrict__ hist, int64_t ntiles) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kTile;
  for (int k = 0; k < kPerThread; ++k) {
    const int64_t i = base + (int64_t)k * kBT + threadIdx.x;
    const bool valid = i < n;
    const uint32_t d = valid ? (uint32_t)(keys[i] >> shift) & 255u : 0u;
    const uint64_t m = match_digit(d, valid);
    if (valid && lane_rank(m) == 0) atomicAdd(&h[d], (uint32_t)__popcll(m));
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of one tile: keys in tile order (step, wave, lane) get
// consecutive slots per digit from the digit's offset for this tile. The
// per-wave digit counts are double-buffered by step parity, so that one
// step needs two barriers.
__global__ __launch_bounds__(kBT) void radix_scatter_kernel(const uint64_t* __restrict__ in,
                                                            uint64_t* __restrict__ out, int64_t n, int shift,
                                                            const uint32_t* __restrict__ goff, int64_t ntiles,
                                                            uint32_t* __restrict__ err) {
  __shared__ uint32_t run[256];
  __shared__ uint32_t wc[2][kWavesBT][256];
  const int tid = (int)threadIdx.x, w = tid >> 6, lane = lane_id();
  run[tid] = goff[(int64_t)tid * ntiles + blockIdx.x];
  const int64_t base = (int64_t)blockIdx.x * kTile;
  for (int k = 0; k < kPerThread; ++k) {
    uint32_t* mine = wc[k & 1][w];
#pragma unroll
    for (int j = 0; j < 256 / kWave; ++j) mine[lane + kWave * j] = 0u;
    const int64_t i = base + (int64_t)k * kBT + tid;
    const bool valid = i < n;
    const uint64_t key = valid ? in[i] : 0ull;
    const uint32_t d = (uint32_t)(key >> shift) & 255u;
    const uint64_t m = match_digit(d, valid);
    const uint32_t rank = lane_rank(m);
    __builtin_amdgcn_wave_barrier();
    if (valid && rank == 0) mine[d] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t pos = 0;
    if (valid) {
      pos = run[d] + rank;
      for (int w2 = 0; w2 < w; ++w2) pos += wc[k & 1][w2][d];
    }
    __syncthreads();
    uint32_t add = 0;
#pragma unroll
    for (int w2 = 0; w2 < kWavesBT; ++w2) add += wc[k & 1][w2][tid];
    run[tid] += add;
    if (valid) {
      if ((int64_t)pos < n)
        out[pos] = key;
      else
        atomicOr(err, 2u);
    }
  }
}

// ---------------------------------------------------------------------------
// Duplicates, degrees, columns
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool first_of_run(const uint64_t* keys, int64_t i, uint64_t k, int vbits, int64_t V) {
  return (int64_t)(k >> vbits) < V && (i == 0 || keys[i - 1] != k);
}

__global__ void unique_flags_kernel(const uint64_t* __restrict__ keys, int64_t n, int vbits, int64_t V,
                                    uint32_t* __restrict__ flag) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    flag[i] = first_of_run(keys, i, keys[i], vbits, V) ? 1u : 0u;
}

// Columns of the distinct keys at their scanned positions, and the degree
// of each source: the keys are sorted, so a wave's distinct keys of one
// source are adjacent lanes and one atomic per source and wave suffices.
// Launched with whole waves over a grid-stride range rounded up to 64.
__global__ void emit_kernel(const uint64_t* __restrict__ keys, int64_t n, int vbits, int64_t V,
                            const uint32_t* __restrict__ pos, int64_t E, int32_t* __restrict__ col,
                            uint32_t* __restrict__ deg, uint32_t* __restrict__ err) {
  const int lane = lane_id();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t nr = (n + kWave - 1) / kWave * kWave;
  const uint64_t vmask = (1ull << vbits) - 1;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nr; i += stride) {
    bool f = false;
    int64_t u = V;
    if (i < n) {
      const uint64_t k = keys[i];
      u = (int64_t)(k >> vbits);
      f = first_of_run(keys, i, k, vbits, V);
      if (f) {
        const uint32_t p = pos[i];
        if ((int64_t)p < E)
          col[p] = (int32_t)(k & vmask);
        else
          atomicOr(err, 4u);
      }
    }
    // A distinct key heads its source's segment when the nearest distinct
    // key in a lower lane has another source (or there is none); the head
    // adds the distinct keys up to the next head. Duplicates never split a
    // source: they sit between keys of that source.
    const uint64_t fm = ballot(f);
    const uint64_t below = fm & ((1ull << lane) - 1ull);
    const int j = below ? 63 - __builtin_clzll(below) : lane;
    const int64_t uj = shfl64(u, j);
    const bool head = f && (below == 0 || uj != u);
    const uint64_t hm = ballot(head);
    if (head) {
      const uint64_t above = lane == kWave - 1 ? 0ull : hm & (~0ull << (lane + 1));
      const uint64_t upto = above ? ((1ull << __builtin_ctzll(above)) - 1ull) : ~0ull;
      const uint64_t seg = fm & upto & (~0ull << lane);
      atomicAdd(&deg[u], (uint32_t)__popcll(seg));
    }
  }
}

__global__ void widen_kernel(const uint32_t* __restrict__ off32, int64_t V, uint32_t E, int64_t* __restrict__ off) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v <= V; v += stride)
    off[v] = v < V ? (int64_t)off32[v] : (int64_t)E;
}

// ---------------------------------------------------------------------------
// Host driver
// ---------------------------------------------------------------------------
static int bits_for(uint64_t x) { return x ? 64 - __builtin_clzll(x) : 0; }

struct BuildMem {
  std::vector<void*> bufs;
  hipStream_t s = nullptr;
  ~BuildMem() {
    if (s) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
    for (void* p : bufs) (void)hipFree(p);
  }
  template <typename T>
  hipError_t alloc(T** p, int64_t n) {
    *p = nullptr;
    hipError_t e = hipMalloc((void**)p, sizeof(T) * (size_t)std::max<int64_t>(n, 1));
    if (e == hipSuccess) bufs.push_back(*p);
    return e;
  }
};

static int grid_for(int64_t n, int cap = 65536) {
  return (int)std::max<int64_t>(1, std::min<int64_t>((n + kBT - 1) / kBT, cap));
}

// Sort the n keys in keys[0] (n keys of `bits` significant bits), build the
// CSR of the distinct valid ones and copy it to out.
static fc_status sort_and_cut(BuildMem& mem, uint64_t* keys0, int64_t n, int vbits, int64_t V, int bits,
                              uint32_t* err, fc_host_graph* out) {
  hipStream_t s = mem.s;
  uint64_t* keys1 = nullptr;
  uint32_t *hist = nullptr, *tmp = nullptr, *flag = nullptr, *deg = nullptr, *dtot = nullptr;
  const int64_t ntiles = tiles_of(n);
  const int64_t hn = 256 * ntiles;
  FC_HIP(mem.alloc(&keys1, n));
  FC_HIP(mem.alloc(&hist, hn));
  FC_HIP(mem.alloc(&tmp, std::max(scan_scratch(hn), std::max(scan_scratch(n), scan_scratch(V)))));
  FC_HIP(mem.alloc(&dtot, 4));
  FC_HIP(hipMemsetAsync(dtot, 0, 4 * sizeof(uint32_t), s));  // the scans of empty ranges write no total
  uint64_t* src = keys0;
  uint64_t* dst = keys1;
  const int passes = n > 0 ? (bits + 7) / 8 : 0;
  for (int pass = 0; pass < passes; ++pass) {
    const int shift = 8 * pass;
    hipLaunchKernelGGL(radix_hist_kernel, dim3((unsigned)ntiles), dim3(kBT), 0, s, src, n, shift, hist, ntiles);
    FC_HIP(hipGetLastError());
    FC_HIP(scan_u32(hist, hist, hn, nullptr, tmp, s));
    hipLaunchKernelGGL(radix_scatter_kernel, dim3((unsigned)ntiles), dim3(kBT), 0, s, src, dst, n, shift,
                       (const uint32_t*)hist, ntiles, err);
    FC_HIP(hipGetLastError());
    std::swap(src, dst);
  }
  // distinct keys
  FC_HIP(mem.alloc(&flag, n));
  hipLaunchKernelGGL(unique_flags_kernel, dim3(grid_for(n)), dim3(kBT), 0, s, src, n, vbits, V, flag);
  FC_HIP(hipGetLastError());
  FC_HIP(scan_u32(flag, flag, n, dtot, tmp, s));
  uint32_t E32 = 0;
  FC_HIP(hipMemcpyAsync(&E32, dtot, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  FC_HIP(hipStreamSynchronize(s));
  const int64_t E = E32;
  int32_t* col = nullptr;
  int64_t* off = nullptr;
  FC_HIP(mem.alloc(&col, E));
  FC_HIP(mem.alloc(&deg, V + 1));
  FC_HIP(mem.alloc(&off, V + 1));
  FC_HIP(hipMemsetAsync(deg, 0, sizeof(uint32_t) * (size_t)(V + 1), s));
  hipLaunchKernelGGL(emit_kernel, dim3(grid_for(n)), dim3(kBT), 0, s, src, n, vbits, V, (const uint32_t*)flag, E, col,
                     deg, err);
  FC_HIP(hipGetLastError());
  FC_HIP(scan_u32(deg, deg, V, dtot + 1, tmp, s));
  hipLaunchKernelGGL(widen_kernel, dim3(grid_for(V + 1)), dim3(kBT), 0, s, (const uint32_t*)deg, V, E32, off);
  FC_HIP(hipGetLastError());
  uint32_t chk[2] = {0, 0};
  FC_HIP(hipMemcpyAsync(chk, dtot + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  FC_HIP(hipMemcpyAsync(chk + 1, err, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  FC_HIP(hipStreamSynchronize(s));
  if (chk[1] & 1u) return fail(FC_ERR_INVALID, "graph build: an endpoint is >= V=%lld", (long long)V);
  if (chk[1] || chk[0] != E32)
    return fail(FC_ERR_HIP, "graph build: inconsistent sort (flags %u, degree sum %u, edges %u)", chk[1], chk[0], E32);
  fc_status st = fc::alloc_host_graph(V, E, out);
  if (st != FC_OK) return st;
  FC_HIP(hipMemcpyAsync(out->offsets, off, sizeof(int64_t) * (size_t)(V + 1), hipMemcpyDeviceToHost, s));
  if (E > 0) FC_HIP(hipMemcpyAsync(out->columns, col, sizeof(int32_t) * (size_t)E, hipMemcpyDeviceToHost, s));
  FC_HIP(hipStreamSynchronize(s));
  out->directed = 0;
  return FC_OK;
}

}  // namespace fcd

using namespace fcd;

extern "C" fc_status fc_csr_from_edges_gpu(const int64_t* edges, int64_t m, int64_t V, fc_host_graph* out) {
  fc::clear_error();
  if (!out || (m > 0 && !edges) || m < 0 || V < 0)
    return fail(FC_ERR_INVALID, "fc_csr_from_edges_gpu: invalid arguments");
  if (m > (int64_t)UINT32_MAX / 2 - kTile)
    return fail(FC_ERR_UNSUPPORTED, "fc_csr_from_edges_gpu: %lld edges exceed 32-bit positions", (long long)m);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(FC_ERR_NOGPU, "no HIP device available");
  BuildMem mem;
  FC_HIP(hipStreamCreateWithFlags(&mem.s, hipStreamNonBlocking));
  int64_t* d_edges = nullptr;
  uint32_t* err = nullptr;
  FC_HIP(mem.alloc(&d_edges, 2 * m));
  FC_HIP(mem.alloc(&err, 1));
  FC_HIP(hipMemsetAsync(err, 0, sizeof(uint32_t), mem.s));
  if (m > 0) FC_HIP(hipMemcpyAsync(d_edges, edges, sizeof(int64_t) * 2 * (size_t)m, hipMemcpyHostToDevice, mem.s));
  if (V == 0) {
    unsigned long long* mx = nullptr;
    FC_HIP(mem.alloc(&mx, 1));
    FC_HIP(hipMemsetAsync(mx, 0, sizeof(unsigned long long), mem.s));
    if (m > 0)
      hipLaunchKernelGGL(max_endpoint_kernel, dim3(grid_for(2 * m, 4096)), dim3(kBT), 0, mem.s, (const int64_t*)d_edges,
                         2 * m, mx);
    FC_HIP(hipGetLastError());
    unsigned long long h = 0;
    FC_HIP(hipMemcpyAsync(&h, mx, sizeof(h), hipMemcpyDeviceToHost, mem.s));
    FC_HIP(hipStreamSynchronize(mem.s));
    V = (int64_t)h;
  }
  if (V > INT32_MAX) return fail(FC_ERR_UNSUPPORTED, "vertex count %lld exceeds int32 ids", (long long)V);
  if (V == 0) {
    fc_status st = fc::alloc_host_graph(0, 0, out);
    if (st == FC_OK) out->offsets[0] = 0;
    return st;
  }
  const int vbits = std::max(1, bits_for((uint64_t)V - 1));
  const int bits = vbits + bits_for((uint64_t)V);
  const int64_t n = 2 * m;
  uint64_t* keys = nullptr;
  FC_HIP(mem.alloc(&keys, n));
  if (m > 0) {
    hipLaunchKernelGGL(edge_keys_kernel, dim3(grid_for(m)), dim3(kBT), 0, mem.s, (const int64_t*)d_edges, m, V, vbits,
                       keys, err);
    FC_HIP(hipGetLastError());
  }
  return sort_and_cut(mem, keys, n, vbits, V, bits, err, out);
}

extern "C" fc_status fc_gen_rmat_gpu(int scale, double edgefactor, double a, double b, double c, uint64_t seed,
                                     fc_host_graph* out) {
  fc::clear_error();
  if (!out || scale < 1 || scale > 30 || edgefactor <= 0)
    return fail(FC_ERR_INVALID, "fc_gen_rmat_gpu: scale must be in [1,30] and edgefactor > 0");
  const double d = 1.0 - a - b - c;
  if (a < 0 || b < 0 || c < 0 || d < -1e-12)
    return fail(FC_ERR_INVALID, "fc_gen_rmat_gpu: quadrant probabilities must be >= 0 and sum to <= 1");
  const int64_t V = 1ll << scale;
  const int64_t M = (int64_t)(edgefactor * (double)V);
  if (M > (int64_t)UINT32_MAX / 2 - kTile)
    return fail(FC_ERR_UNSUPPORTED, "fc_gen_rmat_gpu: %lld edges exceed 32-bit positions", (long long)M);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(FC_ERR_NOGPU, "no HIP device available");
  // the host generator's thresholds and scrambler
  const uint64_t two32 = 1ull << 32;
  const uint64_t ta = (uint64_t)std::llround(a * (double)two32);
  const uint64_t tb = (uint64_t)std::llround((a + b) * (double)two32);
  const uint64_t tc = (uint64_t)std::llround((a + b + c) * (double)two32);
  const fc::Scrambler hp(scale, fc::mix64(seed ^ 0x5CA7ull));
  Scramble perm{hp.mask, hp.m1, hp.m2, hp.a1, hp.a2, hp.shift};
  BuildMem mem;
  FC_HIP(hipStreamCreateWithFlags(&mem.s, hipStreamNonBlocking));
  uint32_t* err = nullptr;
  uint64_t* keys = nullptr;
  const int vbits = std::max(1, bits_for((uint64_t)V - 1));
  const int bits = vbits + bits_for((uint64_t)V);
  FC_HIP(mem.alloc(&err, 1));
  FC_HIP(mem.alloc(&keys, 2 * M));
  FC_HIP(hipMemsetAsync(err, 0, sizeof(uint32_t), mem.s));
  if (M > 0) {
    hipLaunchKernelGGL(rmat_keys_kernel, dim3(grid_for(M)), dim3(kBT), 0, mem.s, scale, M, seed, ta, tb, tc, perm, vbits,
                       keys);
    FC_HIP(hipGetLastError());
  }
  return sort_and_cut(mem, keys, 2 * M, vbits, V, bits, err, out);
}
