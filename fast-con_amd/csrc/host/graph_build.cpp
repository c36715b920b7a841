// Canonical CSR construction, connected components and seeded root sampling.
//
// fc_csr_from_edges gives ST-CON's make_struct semantics
// (st-conn/make_struct.c:304-356 make_g_undirect adds (v,u) for every (u,v)
// and turns self-loops into (-1,-1); st-conn/make_struct.c:553-738 sorts by
// (u,v) and drops adjacent duplicates): the result is a symmetric CSR with
// sorted, duplicate-free adjacency lists and no self-loops. The reference does
// this with Thrust sorts on the device; here it is a parallel counting sort by
// source followed by a per-list sort, which yields the same canonical arrays.
#include "common.hpp"

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <unordered_set>
#include <vector>

#include <omp.h>

using fc::fail;

extern "C" void fc_host_graph_free(fc_host_graph* g) {
  if (!g) return;
  std::free(g->offsets);
  std::free(g->columns);
  g->offsets = nullptr;
  g->columns = nullptr;
  g->V = g->E = 0;
}

namespace fc {

fc_status alloc_host_graph(int64_t V, int64_t E, fc_host_graph* out) {
  std::memset(out, 0, sizeof(*out));
  out->V = V;
  out->E = E;
  out->offsets = (int64_t*)std::malloc(sizeof(int64_t) * (size_t)(V + 1));
  out->columns = (int32_t*)std::malloc(sizeof(int32_t) * (size_t)(E > 0 ? E : 1));
  if (!out->offsets || !out->columns) {
    fc_host_graph_free(out);
    return fail(FC_ERR_NOMEM, "host allocation of a CSR with V=%lld E=%lld failed",
                (long long)V, (long long)E);
  }
  return FC_OK;
}

// Exclusive prefix sum of cnt[0..n) into off[0..n], in parallel blocks.
void exclusive_scan(const int64_t* cnt, int64_t* off, int64_t n) {
  int nt = omp_get_max_threads();
  std::vector<int64_t> part(nt + 1, 0);
#pragma omp parallel num_threads(nt)
  {
    int t = omp_get_thread_num();
    int64_t lo = n * t / nt, hi = n * (t + 1) / nt, s = 0;
    for (int64_t i = lo; i < hi; ++i) s += cnt[i];
    part[t + 1] = s;
#pragma omp barrier
#pragma omp single
    for (int k = 0; k < nt; ++k) part[k + 1] += part[k];
    s = part[t];
    for (int64_t i = lo; i < hi; ++i) {
      off[i] = s;
      s += cnt[i];
    }
  }
  off[n] = part[nt];
}

}  // namespace fc

extern "C" fc_status fc_csr_from_edges(const int64_t* edges, int64_t m, int64_t V,
                                       fc_host_graph* out) {
  fc::clear_error();
  if (!out || (m > 0 && !edges) || m < 0 || V < 0)
    return fail(FC_ERR_INVALID, "fc_csr_from_edges: invalid arguments");
  if (V == 0) {
    int64_t mx = -1;
#pragma omp parallel for reduction(max : mx)
    for (int64_t i = 0; i < 2 * m; ++i) mx = std::max(mx, edges[i]);
    V = mx + 1;
  }
  if (V > INT32_MAX) return fail(FC_ERR_UNSUPPORTED, "vertex count %lld exceeds int32 ids", (long long)V);
  int64_t bad = 0;
#pragma omp parallel for reduction(+ : bad)
  for (int64_t i = 0; i < 2 * m; ++i) bad += edges[i] >= V;
  if (bad) return fail(FC_ERR_INVALID, "fc_csr_from_edges: %lld endpoints >= V=%lld", (long long)bad, (long long)V);

  // 1) degree count over both directions, skipping self-loops and dropped (-1) pairs
  std::vector<int64_t> cnt((size_t)V, 0);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < m; ++i) {
    int64_t u = edges[2 * i], v = edges[2 * i + 1];
    if (u < 0 || v < 0 || u == v) continue;
    __atomic_fetch_add(&cnt[u], 1, __ATOMIC_RELAXED);
    __atomic_fetch_add(&cnt[v], 1, __ATOMIC_RELAXED);
  }
  std::vector<int64_t> raw_off((size_t)V + 1);
  fc::exclusive_scan(cnt.data(), raw_off.data(), V);
  int64_t raw_E = raw_off[V];
  int32_t* raw_col = (int32_t*)std::malloc(sizeof(int32_t) * (size_t)(raw_E > 0 ? raw_E : 1));
  if (!raw_col) return fail(FC_ERR_NOMEM, "host allocation of %lld columns failed", (long long)raw_E);

  // 2) scatter both directions; fill order is irrelevant because lists are sorted next
  std::vector<int64_t> cur(raw_off.begin(), raw_off.end() - 1);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < m; ++i) {
    int64_t u = edges[2 * i], v = edges[2 * i + 1];
    if (u < 0 || v < 0 || u == v) continue;
    raw_col[__atomic_fetch_add(&cur[u], 1, __ATOMIC_RELAXED)] = (int32_t)v;
    raw_col[__atomic_fetch_add(&cur[v], 1, __ATOMIC_RELAXED)] = (int32_t)u;
  }

  // 3) sort and deduplicate each list in place; cnt becomes the unique degree
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t v = 0; v < V; ++v) {
    int32_t* b = raw_col + raw_off[v];
    int32_t* e = raw_col + raw_off[v + 1];
    std::sort(b, e);
    cnt[v] = std::unique(b, e) - b;
  }

  // 4) compact into the output CSR
  std::vector<int64_t> off((size_t)V + 1);
  fc::exclusive_scan(cnt.data(), off.data(), V);
  fc_status st = fc::alloc_host_graph(V, off[V], out);
  if (st != FC_OK) {
    std::free(raw_col);
    return st;
  }
  std::memcpy(out->offsets, off.data(), sizeof(int64_t) * (size_t)(V + 1));
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t v = 0; v < V; ++v)
    std::memcpy(out->columns + off[v], raw_col + raw_off[v], sizeof(int32_t) * (size_t)cnt[v]);
  std::free(raw_col);
  out->directed = 0;
  return FC_OK;
}

extern "C" fc_status fc_components(const fc_host_graph* g, int32_t* label) {
  fc::clear_error();
  if (!g || !label || !g->offsets) return fail(FC_ERR_INVALID, "fc_components: invalid arguments");
  const int64_t V = g->V;
  std::fill(label, label + V, -1);
  std::vector<int32_t> queue((size_t)V);
  // Visiting roots in increasing id order labels each component with its
  // smallest vertex id.
  for (int64_t r = 0; r < V; ++r) {
    if (label[r] >= 0) continue;
    int64_t head = 0, tail = 0;
    queue[tail++] = (int32_t)r;
    label[r] = (int32_t)r;
    while (head < tail) {
      int32_t x = queue[head++];
      for (int64_t e = g->offsets[x]; e < g->offsets[x + 1]; ++e) {
        int32_t y = g->columns[e];
        if (label[y] < 0) {
          label[y] = (int32_t)r;
          queue[tail++] = y;
        }
      }
    }
  }
  return FC_OK;
}

// Acceptance rule of find_stcon_roots (st-conn/reference_functions.c:483-527):
// a candidate is kept when it is not already chosen and has degree > 0 with an
// edge to a vertex other than itself. The candidate stream is this engine's
// own seeded stream, not Graph500's MRG.
extern "C" fc_status fc_sample_roots(const fc_host_graph* g, int32_t n, uint64_t seed, int32_t* out) {
  fc::clear_error();
  if (!g || !out || n < 0) return fail(FC_ERR_INVALID, "fc_sample_roots: invalid arguments");
  const uint64_t V = (uint64_t)g->V;
  if (n == 0) return FC_OK;
  if (V == 0) return fail(FC_ERR_INVALID, "fc_sample_roots: empty graph");
  std::unordered_set<int32_t> chosen;
  chosen.reserve((size_t)n * 2);
  uint64_t tries = 0, limit = 64 * V + 4096;
  int32_t k = 0;
  while (k < n) {
    if (tries >= limit)
      return fail(FC_ERR_INVALID, "fc_sample_roots: only %d of %d roots found", k, n);
    uint64_t w = fc::stream_word(seed, 7, tries++);
    int32_t r = (int32_t)(((unsigned __int128)w * V) >> 64);
    if (chosen.count(r)) continue;
    bool ok = false;
    for (int64_t e = g->offsets[r]; e < g->offsets[r + 1]; ++e)
      if (g->columns[e] != r) {
        ok = true;
        break;
      }
    if (!ok) continue;
    chosen.insert(r);
    out[k++] = r;
  }
  return FC_OK;
}

// Graph500 BFS-tree validation (validate_bfs_result,
// st-conn/reference_functions.c:99-434) for one rank. The depths are derived
// from the parents alone, as the reference's mini-BFS over the predecessor
// map does: here each chain is walked up to a vertex of known depth and the
// depths are filled in on the way back, so the pass is linear; a chain
// longer than V or one that ends outside the root's tree is not a tree.
extern "C" fc_status fc_validate_bfs(const fc_host_graph* g, int32_t root, const int32_t* parents, int64_t nvisited,
                                     int32_t* depth_out) {
  fc::clear_error();
  if (!g || !parents || (g->V > 0 && (!g->offsets || (g->E > 0 && !g->columns))))
    return fail(FC_ERR_INVALID, "fc_validate_bfs: null argument");
  const int64_t V = g->V;
  if (root < 0 || root >= V) return fail(FC_ERR_INVALID, "fc_validate_bfs: root %d out of range", root);
  if (parents[root] != root)
    return fail(FC_ERR_MISMATCH, "Validation error: parent of root vertex %d is %d, not the root itself.", root,
                parents[root]);
  int64_t visited = 0;
  for (int64_t v = 0; v < V; ++v) {
    const int32_t p = parents[v];
    if (v != root && p == v)
      return fail(FC_ERR_MISMATCH, "Validation error: parent of non-root vertex %lld is itself.", (long long)v);
    if (p < -1 || p >= V)
      return fail(FC_ERR_MISMATCH, "Validation error: parent of vertex %lld is out-of-range value %d.", (long long)v, p);
    visited += p != -1;
  }
  if (nvisited >= 0 && visited != nvisited)
    return fail(FC_ERR_MISMATCH, "Validation error: claimed visit count %lld is different from actual count %lld.",
                (long long)nvisited, (long long)visited);
  int64_t bad = -1;
#pragma omp parallel for schedule(dynamic, 4096) reduction(max : bad)
  for (int64_t v = 0; v < V; ++v) {
    const int32_t p = parents[v];
    if (p == -1 || p == v) continue;
    bool found = false;
    for (int64_t e = g->offsets[v]; e < g->offsets[v + 1] && !found; ++e) found = g->columns[e] == p;
    if (!found) bad = std::max(bad, v);
  }
  if (bad >= 0)
    return fail(FC_ERR_MISMATCH, "Validation error: no graph edge from vertex %lld to its parent %d.", (long long)bad,
                parents[bad]);
  // depths from the parents; -2 = not yet known, -1 = outside the tree
  std::vector<int32_t> depth((size_t)V, -2);
  std::vector<int32_t> chain;
  depth[(size_t)root] = 0;
  for (int64_t v = 0; v < V; ++v) {
    if (depth[(size_t)v] != -2) continue;
    if (parents[v] == -1) {
      depth[(size_t)v] = -1;
      continue;
    }
    chain.clear();
    int64_t x = v;
    while (depth[(size_t)x] == -2 && parents[x] != -1 && (int64_t)chain.size() <= V) {
      chain.push_back((int32_t)x);
      x = parents[x];
    }
    const int32_t dx = depth[(size_t)x];
    if (dx < 0 || (int64_t)chain.size() > V)
      return fail(FC_ERR_MISMATCH,
                  "Validation error: BFS predecessors do not form a tree; see vertex %lld and its parent %d.",
                  (long long)v, parents[v]);
    for (size_t k = chain.size(); k-- > 0;) depth[(size_t)chain[k]] = dx + (int32_t)(chain.size() - k);
  }
  int64_t far = -1, out = -1;
#pragma omp parallel for schedule(dynamic, 4096) reduction(max : far, out)
  for (int64_t v = 0; v < V; ++v) {
    const int32_t dv = depth[(size_t)v];
    for (int64_t e = g->offsets[v]; e < g->offsets[v + 1]; ++e) {
      const int32_t dw = depth[(size_t)g->columns[e]];
      if ((dv < 0) != (dw < 0))
        out = std::max(out, v);
      else if (dv >= 0 && (dv - dw > 1 || dw - dv > 1))
        far = std::max(far, v);
    }
  }
  if (out >= 0)
    return fail(FC_ERR_MISMATCH, "Validation error: edge connects vertex %lld to a vertex on the other side of the BFS tree "
                "boundary (depth %d).", (long long)out, depth[(size_t)out]);
  if (far >= 0)
    return fail(FC_ERR_MISMATCH,
                "Validation error: depths of edge endpoints are too far apart (abs. val. > 1) at vertex %lld (depth %d).",
                (long long)far, depth[(size_t)far]);
  if (depth_out) std::copy(depth.begin(), depth.end(), depth_out);
  return FC_OK;
}

namespace fc {

// Structural check of a caller-supplied CSR before it is uploaded: offsets
// start at 0, never decrease and end at E; every column is a vertex id.
// Returns the number of vertices whose degree is >= hub_threshold in *hubs.
template <typename OffT>
static fc_status validate_csr_t(const OffT* off, const int32_t* col, int64_t V, int64_t E, int64_t hub_threshold,
                                int64_t* hubs) {
  if (off[0] != 0 || (int64_t)off[V] != E) return fail(FC_ERR_INVALID, "CSR offsets must start at 0 and end at E");
  int64_t bad_off = 0, bad_col = 0, nh = 0;
#pragma omp parallel for reduction(+ : bad_off, nh)
  for (int64_t v = 0; v < V; ++v) {
    const int64_t d = (int64_t)off[v + 1] - (int64_t)off[v];
    bad_off += d < 0;
    nh += d >= hub_threshold;
  }
  if (bad_off) return fail(FC_ERR_INVALID, "CSR offsets decrease at %lld vertices", (long long)bad_off);
#pragma omp parallel for reduction(+ : bad_col)
  for (int64_t e = 0; e < E; ++e) bad_col += (col[e] < 0 || (int64_t)col[e] >= V);
  if (bad_col) return fail(FC_ERR_INVALID, "%lld CSR columns are not vertex ids in [0,%lld)", (long long)bad_col, (long long)V);
  *hubs = nh;
  return FC_OK;
}

// Device adjacency order: each vertex's neighbours by decreasing degree
// (ties by id), so that a bottom-up scan meets the hubs, which join the
// frontier first, at the front of the list. Levels, verdicts and distances
// do not depend on adjacency order.
template <typename OffT>
static void hub_first_t(const OffT* off, const int32_t* col, int64_t V, int32_t* out) {
  std::vector<int64_t> deg((size_t)V);
#pragma omp parallel for
  for (int64_t v = 0; v < V; ++v) deg[(size_t)v] = (int64_t)off[v + 1] - (int64_t)off[v];
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t v = 0; v < V; ++v) {
    const int64_t b = (int64_t)off[v], e = (int64_t)off[v + 1];
    std::copy(col + b, col + e, out + b);
    std::sort(out + b, out + e, [&](int32_t x, int32_t y) {
      const int64_t dx = deg[(size_t)x], dy = deg[(size_t)y];
      return dx != dy ? dx > dy : x < y;
    });
  }
}

void hub_first32(const int32_t* off, const int32_t* col, int64_t V, int32_t* out) { hub_first_t(off, col, V, out); }
void hub_first64(const int64_t* off, const int32_t* col, int64_t V, int32_t* out) { hub_first_t(off, col, V, out); }

// Degree relabelling: new ids in order of decreasing degree, ties by old id
// (a stable counting sort on degree), so that the per-vertex records of the
// hubs, which most edges lead to, share cache lines. perm[old] = new,
// iperm[new] = old; noff/ncol is the relabelled CSR. Each list is sorted by
// new id unless keep_order, which is the hub-first order above (decreasing
// degree, ties by old id) expressed in new ids.
template <typename OffT>
static void relabel_t(const OffT* off, const int32_t* col, int64_t V, bool keep_order, int32_t* perm,
                      int32_t* iperm, OffT* noff, int32_t* ncol) {
  int64_t maxd = 0;
#pragma omp parallel for reduction(max : maxd)
  for (int64_t v = 0; v < V; ++v) maxd = std::max<int64_t>(maxd, (int64_t)off[v + 1] - (int64_t)off[v]);
  // bucket b holds degree maxd - b, so buckets run by decreasing degree
  std::vector<int64_t> start((size_t)maxd + 2, 0);
  for (int64_t v = 0; v < V; ++v) start[(size_t)(maxd - ((int64_t)off[v + 1] - (int64_t)off[v])) + 1]++;
  for (size_t b = 1; b < start.size(); ++b) start[b] += start[b - 1];
  for (int64_t v = 0; v < V; ++v) {
    const int64_t k = start[(size_t)(maxd - ((int64_t)off[v + 1] - (int64_t)off[v]))]++;
    iperm[k] = (int32_t)v;
  }
#pragma omp parallel for
  for (int64_t k = 0; k < V; ++k) perm[iperm[k]] = (int32_t)k;
  noff[0] = 0;
  for (int64_t k = 0; k < V; ++k) {
    const int32_t v = iperm[k];
    noff[k + 1] = noff[k] + (off[v + 1] - off[v]);
  }
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t k = 0; k < V; ++k) {
    const int32_t v = iperm[k];
    const int64_t b = (int64_t)off[v], e = (int64_t)off[v + 1], nb = (int64_t)noff[k];
    for (int64_t j = b; j < e; ++j) ncol[nb + (j - b)] = perm[col[j]];
    if (!keep_order) std::sort(ncol + nb, ncol + nb + (e - b));
  }
}

void relabel32(const int32_t* off, const int32_t* col, int64_t V, bool keep_order, int32_t* perm, int32_t* iperm,
               int32_t* noff, int32_t* ncol) {
  relabel_t(off, col, V, keep_order, perm, iperm, noff, ncol);
}
void relabel64(const int64_t* off, const int32_t* col, int64_t V, bool keep_order, int32_t* perm, int32_t* iperm,
               int64_t* noff, int32_t* ncol) {
  relabel_t(off, col, V, keep_order, perm, iperm, noff, ncol);
}

fc_status validate_csr32(const int32_t* off, const int32_t* col, int64_t V, int64_t E, int64_t hub, int64_t* hubs) {
  return validate_csr_t(off, col, V, E, hub, hubs);
}
fc_status validate_csr64(const int64_t* off, const int32_t* col, int64_t V, int64_t E, int64_t hub, int64_t* hubs) {
  return validate_csr_t(off, col, V, E, hub, hubs);
}

}  // namespace fc
