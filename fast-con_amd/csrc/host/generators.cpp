// Seeded graph generators: R-MAT, 4-neighbour grid with random edge
// deletion, and Erdős–Rényi G(n, M). Each edge's random words are a pure
// function of (seed, edge index), so results do not depend on thread count.
#include "common.hpp"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

using fc::fail;

namespace fc {
fc_status alloc_host_graph(int64_t V, int64_t E, fc_host_graph* out);
void exclusive_scan(const int64_t* cnt, int64_t* off, int64_t n);
}  // namespace fc

extern "C" fc_status fc_gen_rmat(int scale, double edgefactor, double a, double b, double c,
                                 uint64_t seed, fc_host_graph* out) {
  fc::clear_error();
  if (!out || scale < 1 || scale > 30 || edgefactor <= 0)
    return fail(FC_ERR_INVALID, "fc_gen_rmat: scale must be in [1,30] and edgefactor > 0");
  double d = 1.0 - a - b - c;
  if (a < 0 || b < 0 || c < 0 || d < -1e-12)
    return fail(FC_ERR_INVALID, "fc_gen_rmat: quadrant probabilities must be >= 0 and sum to <= 1");
  const int64_t N = 1ll << scale;
  const int64_t M = (int64_t)(edgefactor * (double)N);
  // Quadrant thresholds on a 32-bit uniform: [0,ta) -> (0,0), [ta,tb) -> (0,1),
  // [tb,tc) -> (1,0), [tc,2^32) -> (1,1).
  const uint64_t two32 = 1ull << 32;
  const uint64_t ta = (uint64_t)std::llround(a * (double)two32);
  const uint64_t tb = (uint64_t)std::llround((a + b) * (double)two32);
  const uint64_t tc = (uint64_t)std::llround((a + b + c) * (double)two32);
  fc::Scrambler perm(scale, fc::mix64(seed ^ 0x5CA7ull));
  int64_t* edges = (int64_t*)std::malloc(sizeof(int64_t) * 2 * (size_t)M);
  if (!edges) return fail(FC_ERR_NOMEM, "host allocation of %lld edge pairs failed", (long long)M);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < M; ++i) {
    uint64_t u = 0, v = 0;
    uint64_t w = 0;
    for (int lvl = 0; lvl < scale; ++lvl) {
      if ((lvl & 1) == 0) w = fc::stream_word(seed, 1, (uint64_t)i * 16 + (uint64_t)(lvl >> 1));
      uint64_t r = (lvl & 1) ? (w >> 32) : (w & 0xFFFFFFFFull);
      uint64_t bu = r >= tb, bv = (r >= ta && r < tb) || r >= tc;
      u = (u << 1) | bu;
      v = (v << 1) | bv;
    }
    edges[2 * i] = (int64_t)perm(u);
    edges[2 * i + 1] = (int64_t)perm(v);
  }
  fc_status st = fc_csr_from_edges(edges, M, N, out);
  std::free(edges);
  return st;
}

extern "C" fc_status fc_gen_er(int64_t n, double mean_degree, uint64_t seed, fc_host_graph* out) {
  fc::clear_error();
  if (!out || n < 1 || n > INT32_MAX || mean_degree < 0)
    return fail(FC_ERR_INVALID, "fc_gen_er: need 1 <= n <= 2^31-1 and mean_degree >= 0");
  const int64_t M = (int64_t)std::llround((double)n * mean_degree / 2.0);
  int64_t* edges = (int64_t*)std::malloc(sizeof(int64_t) * 2 * (size_t)(M > 0 ? M : 1));
  if (!edges) return fail(FC_ERR_NOMEM, "host allocation of %lld edge pairs failed", (long long)M);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < M; ++i) {
    uint64_t w0 = fc::stream_word(seed, 2, 2 * (uint64_t)i);
    uint64_t w1 = fc::stream_word(seed, 2, 2 * (uint64_t)i + 1);
    edges[2 * i] = (int64_t)(((unsigned __int128)w0 * (uint64_t)n) >> 64);
    edges[2 * i + 1] = (int64_t)(((unsigned __int128)w1 * (uint64_t)n) >> 64);
  }
  fc_status st = fc_csr_from_edges(edges, M, n, out);
  std::free(edges);
  return st;
}

// Grid edges are numbered: horizontal (r,c)-(r,c+1) as r*(cols-1)+c, then
// vertical (r,c)-(r+1,c) as rows*(cols-1) + r*cols + c. An edge survives when
// its uniform draw is >= p_delete. The CSR is built directly with ascending
// adjacency (up, left, right, down), identical to canonicalising the edge list.
extern "C" fc_status fc_gen_grid(int32_t rows, int32_t cols, double p_delete, uint64_t seed,
                                 fc_host_graph* out) {
  fc::clear_error();
  if (!out || rows < 1 || cols < 1 || (int64_t)rows * cols > INT32_MAX || p_delete < 0 || p_delete > 1)
    return fail(FC_ERR_INVALID, "fc_gen_grid: invalid size or probability");
  const int64_t V = (int64_t)rows * cols;
  const int64_t H = (int64_t)rows * (cols - 1);
  auto keep_h = [&](int64_t r, int64_t c) {
    return fc::unit_double(fc::stream_word(seed, 3, (uint64_t)(r * (cols - 1) + c))) >= p_delete;
  };
  auto keep_v = [&](int64_t r, int64_t c) {
    return fc::unit_double(fc::stream_word(seed, 3, (uint64_t)(H + r * cols + c))) >= p_delete;
  };
  std::vector<int64_t> cnt((size_t)V);
#pragma omp parallel for schedule(static)
  for (int64_t v = 0; v < V; ++v) {
    int64_t r = v / cols, c = v % cols, k = 0;
    if (r > 0 && keep_v(r - 1, c)) ++k;
    if (c > 0 && keep_h(r, c - 1)) ++k;
    if (c + 1 < cols && keep_h(r, c)) ++k;
    if (r + 1 < rows && keep_v(r, c)) ++k;
    cnt[v] = k;
  }
  std::vector<int64_t> off((size_t)V + 1);
  fc::exclusive_scan(cnt.data(), off.data(), V);
  fc_status st = fc::alloc_host_graph(V, off[V], out);
  if (st != FC_OK) return st;
  std::memcpy(out->offsets, off.data(), sizeof(int64_t) * (size_t)(V + 1));
#pragma omp parallel for schedule(static)
  for (int64_t v = 0; v < V; ++v) {
    int64_t r = v / cols, c = v % cols, p = off[v];
    if (r > 0 && keep_v(r - 1, c)) out->columns[p++] = (int32_t)(v - cols);
    if (c > 0 && keep_h(r, c - 1)) out->columns[p++] = (int32_t)(v - 1);
    if (c + 1 < cols && keep_h(r, c)) out->columns[p++] = (int32_t)(v + 1);
    if (r + 1 < rows && keep_v(r, c)) out->columns[p++] = (int32_t)(v + cols);
  }
  out->directed = 0;
  return FC_OK;
}
