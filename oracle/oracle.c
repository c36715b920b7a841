/*
 * oracle.c — CPU restatement of the reference's host-side algorithms for the
 * BFS / s–t path. TEST INFRASTRUCTURE ONLY: it is the checker used by tests/,
 * by __graft_entry__.smoke() and by bench.py's cpu_baseline leg; the product
 * (libfastcon) never links, loads or calls it.
 *
 * Each function cites the reference file:line it follows (paths relative to
 * the reference checkout; SC = bernaschi/ST_CONN_ONE_ATOMIC/current-st/st-conn).
 * How each one is pinned (golden vectors, closed forms, brute force) is in
 * the header comment of oracle/oracle.py.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <time.h>

/* ---------------------------------------------------------------------- */
/* Graph::BfsInit + Graph::bfs (BFS-4K/src/host/graph.cpp:70-97), with the  */
/* statistics of visitedNodes / visitedEdges / getMaxDistance (:99-114).    */
/* inf is the value written for unreached vertices (the reference uses      */
/* INF = (dist_t)-1 = 65535, BFS-4K/config.h:14,50).                        */
/* queue_out (nullable, V entries) receives the FIFO order.                 */
/* info[0] = visitedNodes, info[1] = visitedEdges, info[2] = getMaxDistance */
/* ---------------------------------------------------------------------- */
int orc_bfs(int64_t V, int64_t E, const int64_t* off, const int32_t* col, int32_t src, int32_t inf,
            int32_t* dist, int32_t* queue_out, int64_t* info) {
  if (src < 0 || src >= V) return -1;
  unsigned char* visited = (unsigned char*)calloc((size_t)V, 1);
  int32_t* queue = queue_out ? queue_out : (int32_t*)malloc(sizeof(int32_t) * (size_t)V);
  if (!visited || !queue) return -2;
  int64_t left = 0, right = 1;
  for (int64_t i = 0; i < V; ++i) dist[i] = inf;
  visited[src] = 1;
  dist[src] = 0;
  queue[0] = src;
  while (left < right) {
    int32_t q = queue[left++];
    for (int64_t i = off[q]; i < off[q + 1]; ++i) {
      int32_t d = col[i];
      if (!visited[d]) {
        visited[d] = 1;
        dist[d] = dist[q] + 1;
        queue[right++] = d;
      }
    }
  }
  if (info) {
    info[0] = right;
    /* visitedEdges: all E when every vertex was reached, else the sum of the
     * out-degrees of the dequeued vertices (graph.cpp:103-110). */
    int64_t sum = 0;
    if (right == V) {
      sum = E;
    } else {
      for (int64_t i = 0; i < right; ++i) sum += off[queue[i] + 1] - off[queue[i]];
    }
    info[1] = sum;
    info[2] = dist[queue[right - 1]];
  }
  free(visited);
  if (!queue_out) free(queue);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* Graph::ToCSR (BFS-4K/src/host/graph.cpp:19-42): counting sort of a COO   */
/* list; adjacency in COO order; undirected adds the reverse entry right    */
/* after. No self-loop or duplicate removal. off: V+1, col: E entries.      */
/* ---------------------------------------------------------------------- */
int orc_to_csr(int64_t V, const int32_t* coo, int64_t n, int undirected, int64_t* off, int32_t* col) {
  int64_t* deg = (int64_t*)calloc((size_t)V, sizeof(int64_t));
  int64_t* tmp = (int64_t*)calloc((size_t)V, sizeof(int64_t));
  if (!deg || !tmp) return -2;
  for (int64_t i = 0; i < n; ++i) {
    int32_t a = coo[2 * i], b = coo[2 * i + 1];
    if (a < 0 || a >= V || b < 0 || b >= V) {
      free(deg);
      free(tmp);
      return -1;
    }
    deg[a]++;
    if (undirected) deg[b]++;
  }
  off[0] = 0;
  for (int64_t v = 0; v < V; ++v) off[v + 1] = off[v] + deg[v];
  for (int64_t i = 0; i < n; ++i) {
    int32_t a = coo[2 * i], b = coo[2 * i + 1];
    col[off[a] + tmp[a]++] = b;
    if (undirected) col[off[b] + tmp[b]++] = a;
  }
  free(deg);
  free(tmp);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* ST-CON canonicalisation (SC/make_struct.c:304-356 make_g_undirect: add   */
/* (v,u) for each (u,v), self-loops become (-1,-1); SC/make_struct.c:553-738 */
/* make_own_csr_nomulti: stable sort by (u,v), drop adjacent duplicates).   */
/* Restated as: collect both directions, drop negatives and self-loops,     */
/* qsort the pairs, skip repeats. Returns E; off: V+1, col: capacity 2m.    */
/* ---------------------------------------------------------------------- */
static int cmp_pair(const void* a, const void* b) {
  const int64_t* x = (const int64_t*)a;
  const int64_t* y = (const int64_t*)b;
  if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
  if (x[1] != y[1]) return x[1] < y[1] ? -1 : 1;
  return 0;
}

int64_t orc_canonical(const int64_t* edges, int64_t m, int64_t V, int64_t* off, int32_t* col) {
  int64_t* pr = (int64_t*)malloc(sizeof(int64_t) * 4 * (size_t)(m > 0 ? m : 1));
  if (!pr) return -2;
  int64_t k = 0;
  for (int64_t i = 0; i < m; ++i) {
    int64_t u = edges[2 * i], v = edges[2 * i + 1];
    if (u < 0 || v < 0 || u == v) continue;
    if (u >= V || v >= V) {
      free(pr);
      return -1;
    }
    pr[2 * k] = u;
    pr[2 * k + 1] = v;
    ++k;
    pr[2 * k] = v;
    pr[2 * k + 1] = u;
    ++k;
  }
  qsort(pr, (size_t)k, 2 * sizeof(int64_t), cmp_pair);
  int64_t E = 0;
  for (int64_t v = 0; v <= V; ++v) off[v] = 0;
  for (int64_t i = 0; i < k; ++i) {
    if (i > 0 && pr[2 * i] == pr[2 * i - 2] && pr[2 * i + 1] == pr[2 * i - 1]) continue;
    col[E++] = (int32_t)pr[2 * i + 1];
    off[pr[2 * i] + 1]++;
  }
  for (int64_t v = 0; v < V; ++v) off[v + 1] += off[v];
  free(pr);
  return E;
}

/* ---------------------------------------------------------------------- */
/* s–t on the host: BFS from s, read the level of t — the reference's own   */
/* oracle path (Graph::bfs, graph.cpp:83-97). ST-CON's DEPTH is the meeting */
/* iteration of its two-colour visit, ceil(d/2) (SC/make_stcon.c:971,1496). */
/* ---------------------------------------------------------------------- */
int orc_stcon(int64_t V, int64_t E, const int64_t* off, const int32_t* col, int32_t s, int32_t t,
              int32_t* connected, int32_t* dist, int32_t* depth) {
  int32_t* d = (int32_t*)malloc(sizeof(int32_t) * (size_t)V);
  if (!d) return -2;
  int rc = orc_bfs(V, E, off, col, s, -1, d, NULL, NULL);
  if (rc) {
    free(d);
    return rc;
  }
  *dist = d[t];
  *connected = d[t] >= 0;
  *depth = d[t] >= 0 ? (d[t] + 1) / 2 : -1;
  free(d);
  return 0;
}

/* Work model of one plain two-colour s–t query: the vertices the visit
 * expands are those within D-1 of s or of t, D = ceil(d/2); if s and t are
 * not connected, both components. Returns |V_x| and E_x = sum of degrees. */
int orc_stcon_work(int64_t V, int64_t E, const int64_t* off, const int32_t* col, int32_t s, int32_t t,
                   int64_t* vx, int64_t* ex) {
  int32_t* ds = (int32_t*)malloc(sizeof(int32_t) * (size_t)V);
  int32_t* dt = (int32_t*)malloc(sizeof(int32_t) * (size_t)V);
  if (!ds || !dt) return -2;
  orc_bfs(V, E, off, col, s, -1, ds, NULL, NULL);
  orc_bfs(V, E, off, col, t, -1, dt, NULL, NULL);
  int32_t d = ds[t];
  int64_t nv = 0, ne = 0;
  for (int64_t v = 0; v < V; ++v) {
    int in;
    if (d < 0)
      in = ds[v] >= 0 || dt[v] >= 0;
    else {
      int32_t D = (d + 1) / 2;
      in = (ds[v] >= 0 && ds[v] <= D - 1) || (dt[v] >= 0 && dt[v] <= D - 1);
    }
    if (in) {
      nv++;
      ne += off[v + 1] - off[v];
    }
  }
  *vx = nv;
  *ex = ne;
  free(ds);
  free(dt);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* BFS-tree validity, the rules of the Graph500 validator                  */
/* (SC/reference_functions.c:108-425) for one BFS on an undirected graph:  */
/*  1 the root is its own parent at level 0;                               */
/*  2 every reached vertex other than the root has a parent whose level is */
/*    one less and the tree edge exists in the graph;                      */
/*  3 every edge joins vertices whose levels differ by at most one, and    */
/*    either both or neither endpoint is reached;                          */
/*  4 unreached vertices have no parent.                                   */
/* Returns 0 when valid, else the number of the first rule violated.       */
/* ---------------------------------------------------------------------- */
int orc_validate_bfs_tree(int64_t V, const int64_t* off, const int32_t* col, int32_t root, const int32_t* level,
                          const int32_t* parent) {
  if (parent[root] != root || level[root] != 0) return 1;
  for (int64_t v = 0; v < V; ++v) {
    if (level[v] < 0) {
      if (parent[v] != -1) return 4;
      continue;
    }
    if (v == root) continue;
    int32_t p = parent[v];
    if (p < 0 || p >= V || level[p] != level[v] - 1) return 2;
    int found = 0;
    for (int64_t e = off[p]; e < off[p + 1]; ++e)
      if (col[e] == v) {
        found = 1;
        break;
      }
    if (!found) return 2;
  }
  for (int64_t v = 0; v < V; ++v)
    for (int64_t e = off[v]; e < off[v + 1]; ++e) {
      int32_t u = col[e];
      if ((level[v] < 0) != (level[u] < 0)) return 3;
      if (level[v] >= 0 && abs(level[v] - level[u]) > 1) return 3;
    }
  return 0;
}

/* count_visited_edges (SC/adj_func.c:334-357): for every reached vertex,
 * count its edges whose target id is <= its own id (each undirected edge once). */
int64_t orc_count_visited_edges(int64_t V, const int64_t* off, const int32_t* col, const int32_t* level) {
  int64_t c = 0;
  for (int64_t v = 0; v < V; ++v)
    if (level[v] >= 0)
      for (int64_t e = off[v]; e < off[v + 1]; ++e)
        if (col[e] <= v) ++c;
  return c;
}

/* get_statistics (SC/reference_functions.c:531-557): mean, std (n-1), and
 * order statistics. r = {min, firstquartile, median, thirdquartile, max, mean, std}. */
static int cmp_double(const void* a, const void* b) {
  double x = *(const double*)a, y = *(const double*)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

void orc_get_statistics(const double* x, int n, double* r) {
  double t = 0;
  for (int i = 0; i < n; ++i) t += x[i];
  t /= n;
  r[5] = t;
  double s = 0;
  for (int i = 0; i < n; ++i) s += (x[i] - r[5]) * (x[i] - r[5]);
  s /= n - 1;
  r[6] = sqrt(s);
  double* xx = (double*)malloc(sizeof(double) * (size_t)n);
  memcpy(xx, x, sizeof(double) * (size_t)n);
  qsort(xx, (size_t)n, sizeof(double), cmp_double);
  r[0] = xx[0];
  r[1] = (xx[(n - 1) / 4] + xx[n / 4]) * .5;
  r[2] = (xx[(n - 1) / 2] + xx[n / 2]) * .5;
  r[3] = (xx[n - 1 - (n - 1) / 4] + xx[n - 1 - n / 4]) * .5;
  r[4] = xx[n - 1];
  free(xx);
}

/* ---------------------------------------------------------------------- */
/* find_stcon_roots (SC/reference_functions.c:483-527) on one rank: draw    */
/* r = (int64)((d0+d1)*nglobalverts) mod nglobalverts from the random-access */
/* stream at position counter (+2 per try); skip repeats; accept when r has */
/* degree > 0 and an edge to a vertex other than itself. rng is Graph500's  */
/* make_random_numbers (signature of GEN/make_graph.h:38-44).               */
/* ---------------------------------------------------------------------- */
typedef void (*orc_rng_fn)(int64_t nvalues, uint64_t seed1, uint64_t seed2, int64_t position, double* result);

int orc_find_stcon_roots(int n, int64_t nglobalverts, const int64_t* off, const int32_t* col, uint64_t seed1,
                         uint64_t seed2, orc_rng_fn rng, int64_t* roots) {
  uint64_t counter = 0;
  int idx;
  for (idx = 0; idx < n; ++idx) {
    int64_t root;
    while (1) {
      double d[2];
      rng(2, seed1, seed2, (int64_t)counter, d);
      root = (int64_t)((d[0] + d[1]) * (double)nglobalverts) % nglobalverts;
      counter += 2;
      if (counter > 2 * (uint64_t)nglobalverts) break;
      int dup = 0;
      for (int i = 0; i < idx; ++i)
        if (roots[i] == root) {
          dup = 1;
          break;
        }
      if (dup) continue;
      int ok = 0;
      for (int64_t e = off[root]; e < off[root + 1]; ++e)
        if (col[e] != root && off[root + 1] - off[root] > 0) {
          ok = 1;
          break;
        }
      if (ok) break;
    }
    roots[idx] = root;
  }
  return idx;
}

/* ---------------------------------------------------------------------- */
/* Batched bit-parallel s–t, restated with per-query bit sets: wave of up   */
/* to 64 queries, both sides advanced one level per iteration. Summed over  */
/* iterations and sides it reports A (frontier vertices whose word has a    */
/* live query) and their degree sum, and returns per-query distances. This  */
/* is the work model of the batched engine's plain schedule.                */
/* ---------------------------------------------------------------------- */
int orc_batch_wave(int64_t V, const int64_t* off, const int32_t* col, int nq, const int32_t* s, const int32_t* t,
                   int32_t* dist, int64_t* a_sum, int64_t* e_sum, int32_t* iters) {
  uint64_t *vs = calloc((size_t)V, 8), *vt = calloc((size_t)V, 8), *fs = calloc((size_t)V, 8),
           *ft = calloc((size_t)V, 8), *ns = calloc((size_t)V, 8), *nt = calloc((size_t)V, 8);
  if (!vs || !vt || !fs || !ft || !ns || !nt) return -2;
  uint64_t active = 0;
  for (int q = 0; q < nq; ++q) {
    uint64_t b = 1ull << q;
    dist[q] = -1;
    if (s[q] == t[q]) {
      dist[q] = 0;
      continue;
    }
    active |= b;
    vs[s[q]] |= b;
    fs[s[q]] |= b;
    vt[t[q]] |= b;
    ft[t[q]] |= b;
  }
  int64_t A = 0, Es = 0;
  int k = 0;
  while (active) {
    ++k;
    for (int64_t v = 0; v < V; ++v) {
      uint64_t a = (fs[v] | ft[v]) & active;
      if (!a) continue;
      /* work counted per side: a vertex in both frontiers is expanded twice */
      int sides = ((fs[v] & active) != 0) + ((ft[v] & active) != 0);
      A += sides;
      Es += sides * (off[v + 1] - off[v]);
      for (int64_t e = off[v]; e < off[v + 1]; ++e) {
        int32_t u = col[e];
        ns[u] |= fs[v] & active;
        nt[u] |= ft[v] & active;
      }
    }
    uint64_t odd = 0, even = 0, anys = 0, anyt = 0;
    for (int64_t v = 0; v < V; ++v) {
      uint64_t xs = ns[v] & ~vs[v], xt = nt[v] & ~vt[v];
      odd |= (xs & vt[v]) | (xt & vs[v]);
      even |= xs & xt;
      vs[v] |= xs;
      vt[v] |= xt;
      fs[v] = xs;
      ft[v] = xt;
      ns[v] = nt[v] = 0;
      anys |= xs;
      anyt |= xt;
    }
    for (int q = 0; q < nq; ++q) {
      uint64_t b = 1ull << q;
      if (!(active & b)) continue;
      if (odd & b)
        dist[q] = 2 * k - 1;
      else if (even & b)
        dist[q] = 2 * k;
      else if (!(anys & b) || !(anyt & b))
        dist[q] = -1;
      else
        continue;
      active &= ~b;
    }
  }
  *a_sum = A;
  *e_sum = Es;
  *iters = k;
  free(vs);
  free(vt);
  free(fs);
  free(ft);
  free(ns);
  free(nt);
  return 0;
}

/* CPU baseline: n host BFS s–t queries (BFS from s, read level of t),
 * single thread, timed with CLOCK_MONOTONIC. Returns seconds. */
double orc_time_stcon(int64_t V, int64_t E, const int64_t* off, const int32_t* col, int n, const int32_t* s,
                      const int32_t* t, int32_t* dist) {
  int32_t* d = (int32_t*)malloc(sizeof(int32_t) * (size_t)V);
  int32_t* q = (int32_t*)malloc(sizeof(int32_t) * (size_t)V);
  struct timespec a, b;
  clock_gettime(CLOCK_MONOTONIC, &a);
  for (int i = 0; i < n; ++i) {
    orc_bfs(V, E, off, col, s[i], -1, d, q, NULL);
    dist[i] = d[t[i]];
  }
  clock_gettime(CLOCK_MONOTONIC, &b);
  free(d);
  free(q);
  return (double)(b.tv_sec - a.tv_sec) + 1e-9 * (double)(b.tv_nsec - a.tv_nsec);
}

/* CPU baseline for BFS: n host BFSs; returns seconds, sums visitedEdges. */
double orc_time_bfs(int64_t V, int64_t E, const int64_t* off, const int32_t* col, int n, const int32_t* roots,
                    int64_t* edges_total) {
  int32_t* d = (int32_t*)malloc(sizeof(int32_t) * (size_t)V);
  int32_t* q = (int32_t*)malloc(sizeof(int32_t) * (size_t)V);
  int64_t info[3], tot = 0;
  struct timespec a, b;
  clock_gettime(CLOCK_MONOTONIC, &a);
  for (int i = 0; i < n; ++i) {
    orc_bfs(V, E, off, col, roots[i], -1, d, q, info);
    tot += info[1];
  }
  clock_gettime(CLOCK_MONOTONIC, &b);
  *edges_total = tot;
  free(d);
  free(q);
  return (double)(b.tv_sec - a.tv_sec) + 1e-9 * (double)(b.tv_nsec - a.tv_nsec);
}
