/*
 * fastcon.h — C ABI of libfastcon, the MI355X (gfx950) BFS and s–t
 * connectivity engine.
 *
 * Plain C, plain pointers and sizes; no HIP or CUDA types cross this line.
 * Every entry point returns an fc_status and never calls exit(); the text of
 * the last error on the calling thread is available from fc_last_error().
 *
 * Reference interfaces replaced (paths relative to the reference checkout):
 *   BFS-4K   class cudaGraph {cudaGraph(Graph&), Reset, cudaBFS4K, cudaBFS4K_N}
 *            BFS-4K/include/device/cudaGraph.cuh:14-35
 *   BFS-4K   class Graph (host CSR, host BFS oracle)
 *            BFS-4K/include/host/graph.h:15-42, BFS-4K/src/host/graph.cpp:19-114
 *   BFS-4K   readGraph::readGraphHeader / readSTD
 *            BFS-4K/Util/readGraph.h:17-23
 *   BFS_4K_DS double-source visit (found_common)
 *            BFS_4K_DS/src/device/BFS_WorkEfficient.cu:8-98, BFS_4K_DS/src/device/cudaGraph.cu:5
 *   ST-CON   extern "C" make_stcon(...)
 *            bernaschi/ST_CONN_ONE_ATOMIC/current-st/st-conn/include/make_stcon.h:11-15
 *   ST-CON   run_make_struct (symmetrise, drop self-loops and duplicates)
 *            bernaschi/ST_CONN_ONE_ATOMIC/current-st/st-conn/make_struct.c:304-356,553-738
 *   ST-CON   find_stcon_roots (seeded root sampling)
 *            bernaschi/ST_CONN_ONE_ATOMIC/current-st/st-conn/reference_functions.c:483-527
 *
 * Encodings:
 *   FC_INF (-1) marks an unreached vertex in level arrays. The reference
 *   writes INF = (dist_t)-1 = 65535 (BFS-4K/config.h:14,50) into an int array;
 *   a level of 65535 never occurs on the supported inputs, so the two
 *   encodings map one to one (65535 <-> -1).
 *   Parents: -1 for unreached vertices, the root points to itself (as the
 *   ST-CON predecessor reconstruction does, st-conn/make_stcon.c:1446-1455).
 *
 * Threading: one fc_graph per host thread; each fc_graph owns one HIP stream
 * (default priority) and its own device workspace. Calls are synchronous from
 * the caller's point of view.
 */
#ifndef FASTCON_H
#define FASTCON_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FC_INF (-1)
#define FC_VERSION_MAJOR 0
#define FC_VERSION_MINOR 1

typedef enum fc_status {
  FC_OK = 0,
  FC_ERR_INVALID = 1,   /* bad argument (null pointer, vertex out of range, ...) */
  FC_ERR_HIP = 2,       /* HIP runtime error */
  FC_ERR_NOMEM = 3,     /* host or device allocation failed */
  FC_ERR_IO = 4,        /* unreadable or non-regular file (BFS-4K/Util/readUtil.cpp:52-61) */
  FC_ERR_FORMAT = 5,    /* graph format not recognised (BFS-4K/Util/readGraphHeader.cpp:37-38) */
  FC_ERR_CAPACITY = 6,  /* a fixed-capacity buffer would overflow (BFS-4K/src/device/BFS_WorkEfficient.cu:111-112) */
  FC_ERR_NOGPU = 7,     /* no HIP device, or the device code object is missing */
  FC_ERR_TIMEOUT = 8,   /* a device-side bounded wait expired; results are invalid */
  FC_ERR_UNSUPPORTED = 9,
  FC_ERR_MISMATCH = 10  /* a validator found a wrong result (validate_bfs_result, validate_stcon_bfs) */
} fc_status;

/* Text of the last error raised on the calling thread ("" if none). */
const char* fc_last_error(void);
/* "fastcon <major>.<minor> gfx950" */
const char* fc_version(void);

/* ------------------------------------------------------------------------ */
/* Host graphs (CSR in host memory). Built by the readers, the generators and */
/* the canonicaliser below; released with fc_host_graph_free.                */
/* ------------------------------------------------------------------------ */

typedef struct fc_host_graph {
  int64_t  V;          /* vertices */
  int64_t  E;          /* directed edges = length of columns */
  int64_t* offsets;    /* V+1 entries; adjacency of v is columns[offsets[v] .. offsets[v+1]) */
  int32_t* columns;    /* E entries */
  int32_t  directed;   /* 0: every edge is stored in both directions */
  int32_t  reserved;
} fc_host_graph;

/* Direction selector for the file readers: as BFS-4K's -D / -U switches. */
typedef enum fc_direction {
  FC_DIR_FILE = 0,        /* take it from the file header, directed when undefined */
  FC_DIR_DIRECTED = 1,    /* -D */
  FC_DIR_UNDIRECTED = 2   /* -U */
} fc_direction;

void fc_host_graph_free(fc_host_graph* g);

/* Read a graph file in any FileFormat.txt format (MatrixMarket, DIMACS9 /
 * GTgraph, DIMACS10 / METIS, SNAP; PGSolver headers are recognised but have
 * no body reader, as in the reference). Vertex numbering, direction rules and
 * adjacency order follow readGraphHeader + readSTD + Graph::ToCSR
 * (BFS-4K/Util/readGraphHeader.cpp:14-114, BFS-4K/Util/readGraphSTD.cpp:14-145,
 * BFS-4K/src/host/graph.cpp:19-66): no deduplication, adjacency in file order. */
fc_status fc_read_graph(const char* path, fc_direction dir, fc_host_graph* out);

/* Read the ST-CON edge-list loader's format (st-conn/load_graph.c:20-104). */
fc_status fc_read_edge_list(const char* path, int64_t** edges_out, int64_t* nedges_out,
                            int64_t* nverts_out);
void fc_free(void* p);

/* Canonical undirected CSR from an edge list, with the semantics of ST-CON's
 * make_struct (st-conn/make_struct.c:304-356,553-738): every (u,v) also adds
 * (v,u); self-loops and edges with a negative endpoint are dropped; duplicate
 * pairs are removed; each adjacency list is sorted ascending.
 * edges: 2*m int64 values (u0,v0,u1,v1,...). V: vertex count (0 = max id + 1). */
fc_status fc_csr_from_edges(const int64_t* edges, int64_t m, int64_t V, fc_host_graph* out);

/* Seeded generators (this engine's own; deterministic for a seed at any
 * thread count). All return canonical undirected CSR.
 * R-MAT: 2^scale vertices, edgefactor*2^scale generated pairs, quadrant
 * probabilities a, b, c and d = 1-a-b-c, vertex ids scrambled by a seeded
 * bijection of [0, 2^scale). */
fc_status fc_gen_rmat(int scale, double edgefactor, double a, double b, double c,
                      uint64_t seed, fc_host_graph* out);
/* rows x cols 4-neighbour grid, vertex r*cols+c; each undirected edge is
 * deleted with probability p_delete. */
fc_status fc_gen_grid(int32_t rows, int32_t cols, double p_delete, uint64_t seed,
                      fc_host_graph* out);
/* Erdős–Rényi G(n, M) with M = n*mean_degree/2 uniformly drawn pairs. */
fc_status fc_gen_er(int64_t n, double mean_degree, uint64_t seed, fc_host_graph* out);

/* n distinct seeded vertices of degree > 0 (the acceptance rule of
 * find_stcon_roots, st-conn/reference_functions.c:483-527, drawn from this
 * engine's own stream). Returns FC_ERR_INVALID if fewer exist. */
fc_status fc_sample_roots(const fc_host_graph* g, int32_t n, uint64_t seed, int32_t* out);

/* The same two builders on the GPU, with the same results bit for bit: the
 * edges are generated on the device (R-MAT) or uploaded, and the canonical
 * CSR is built there by an in-house LSD radix sort on (u, v) keys, duplicate
 * removal and scans, then copied back. Replaces make_struct's Thrust sorts
 * and unique (st-conn/make_struct_gpufunc.cu:242-328,
 * st-conn/cudakernel/gstruct_kernel.cu:114-145,187-235). Limits: 2*m below
 * 2^32 - 2^13 (E fits 32-bit positions); needs about 20 bytes of device
 * memory per directed edge. */
fc_status fc_gen_rmat_gpu(int scale, double edgefactor, double a, double b, double c,
                          uint64_t seed, fc_host_graph* out);
fc_status fc_csr_from_edges_gpu(const int64_t* edges, int64_t m, int64_t V, fc_host_graph* out);

/* Connected-component labels (label = smallest vertex id of the component). */
fc_status fc_components(const fc_host_graph* g, int32_t* label_out);

/* Graph500 BFS-tree validation on the host graph (validate_bfs_result,
 * st-conn/reference_functions.c:99-434), one rank: the root is its own
 * parent and no other vertex is; parents are -1 or vertex ids; nvisited
 * (ignored when < 0) equals the number of vertices with a parent; each
 * parent is a neighbour; the parents form a tree (depths follow from the
 * root); every edge joins two tree vertices whose depths differ by at most
 * one, or two vertices outside the tree. depth_out (nullable, V entries)
 * receives the tree depths, -1 outside the tree. FC_OK, or FC_ERR_MISMATCH
 * with the first violation in the reference's wording. */
fc_status fc_validate_bfs(const fc_host_graph* g, int32_t root, const int32_t* parents, int64_t nvisited,
                          int32_t* depth_out);

/* ------------------------------------------------------------------------ */
/* Device graphs and queries.                                                */
/* ------------------------------------------------------------------------ */

typedef struct fc_graph fc_graph;

/* Upload a CSR graph; the library keeps device copies and allocates all
 * per-query workspace once. Replaces cudaGraph::cudaGraph
 * (BFS-4K/src/device/cudaGraph.cu:7-36). By default the device copy
 * numbers the vertices by decreasing degree (ties by id), so that the
 * per-vertex state of the hubs shares cache lines, and lists each vertex's
 * neighbours by decreasing degree, which makes bottom-up levels stop early.
 * Every call takes and returns the caller's vertex ids; levels, verdicts and
 * distances do not depend on either choice. flags: FC_GRAPH_KEEP_ORDER keeps
 * the caller's adjacency order (then BFS parents follow it);
 * FC_GRAPH_KEEP_IDS keeps the caller's numbering on the device. */
#define FC_GRAPH_KEEP_ORDER 1u
#define FC_GRAPH_KEEP_IDS 2u
fc_status fc_graph_create_csr32(const int32_t* off, const int32_t* col, int32_t V, int64_t E,
                                uint32_t flags, fc_graph** out);
fc_status fc_graph_create_csr64(const int64_t* off, const int32_t* col, int64_t V, int64_t E,
                                uint32_t flags, fc_graph** out);
fc_status fc_graph_destroy(fc_graph* g);

/* Per-call statistics. */
typedef struct fc_stats {
  double  time_ms;            /* device time of the visit (HIP events on the graph's stream) */
  int32_t levels;             /* BFS: number of non-empty levels; s–t: visit iterations */
  int32_t depth;              /* s–t: iteration at which the frontiers met (ST-CON DEPTH) */
  int64_t visited_vertices;   /* BFS: reached vertices V_r; s–t: labelled vertices */
  int64_t visited_edges;      /* BFS: sum of degrees of reached vertices E_r */
  int64_t expanded_vertices;  /* frontier vertices whose adjacency was scanned */
  int64_t expanded_edges;     /* sum of degrees of the expanded vertices */
  int32_t topdown_levels;
  int32_t bottomup_levels;
  int32_t max_frontier;
  int32_t launches;           /* working kernel launches of the visit (each query's reset included) */
} fc_stats;

/* Single-source BFS. Replaces Reset + cudaBFS4K + readback
 * (BFS-4K/src/device/BFS_WorkEfficient.cu:81-133).
 * levels_out / parents_out: host arrays of V entries, or NULL to leave the
 * result in device memory (fetch later with fc_graph_levels / fc_graph_parents).
 * Levels equal Graph::bfs Distance[] (BFS-4K/src/host/graph.cpp:83-97). */
fc_status fc_bfs(fc_graph* g, int32_t src, int32_t* levels_out, int32_t* parents_out,
                 fc_stats* stats);
fc_status fc_graph_levels(fc_graph* g, int32_t* levels_out);
fc_status fc_graph_parents(fc_graph* g, int32_t* parents_out);

/* Plain s–t connectivity: one query per launch, two colours, one atomicCAS
 * per vertex, both frontiers expanded each iteration, stop when they meet
 * (ST-CON make_stcon, st-conn/make_stcon.c:730-1497, and the BFS_4K_DS
 * visit, BFS_4K_DS/src/device/WorkEfficientKernel/BFS_WE_Kernels1.cu:25-89).
 * connected: 1/0. dist: exact hop distance (Graph::bfs level of t from s),
 * -1 if not connected. meet: a vertex on a shortest path where the two
 * colours met, or -1 (nullable). stats->depth is ST-CON's DEPTH = ceil(dist/2). */
fc_status fc_stcon(fc_graph* g, int32_t s, int32_t t, int32_t* connected, int32_t* dist,
                   int32_t* meet, fc_stats* stats);

/* Many plain queries, each still one launch (the reference's per-pair loop,
 * st-conn/make_stcon.c:418-452). */
fc_status fc_stcon_many(fc_graph* g, const int32_t* s, const int32_t* t, int32_t n,
                        int32_t* connected, int32_t* dist, fc_stats* stats);

/* fc_stcon_many without any device allocation. No counterpart in the
 * reference (whose ST-CON wrappers cudaMalloc and cudaFree scratch on every
 * call); it is for callers that cannot allocate device memory, such as
 * allocation-free serving loops and harnesses that own all device memory.
 * d_off (V+1 int32) and d_col (E int32) are a CSR graph in the caller's
 * device memory, used in the caller's numbering and adjacency order; s, t,
 * connected, dist and stats are host memory as in fc_stcon_many, with the
 * same verdicts and distances. The per-query workspace is one fixed device
 * array of about 27 MB that every GPU the library runs on holds for the life
 * of the process. Limits: 1 <= V <= 2^20 and E <= 2^25 (FC_ERR_INVALID
 * beyond them). The call creates no stream, event or graph; it launches on
 * the null stream and returns when done. Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU race and must be serialised.
 * stats->time_ms is host wall time. */
fc_status fc_stcon_many_static(const int32_t* d_off, const int32_t* d_col, int32_t V, int64_t E,
                               const int32_t* s, const int32_t* t, int32_t n, int32_t* connected,
                               int32_t* dist, fc_stats* stats);

/* One plain query with a shortest path, in caller ids: path_out[0] = s,
 * path_out[dist] = t and *len_out = dist + 1 (0 when not connected, 1 when
 * s == t). cap must be at least dist + 1, else FC_ERR_CAPACITY with
 * *len_out set to the length needed. The visit records each claimed
 * vertex's parent; the path joins the s-side and t-side parent chains at the
 * meeting edge. Replaces the predecessor reconstruction of make_stcon
 * (st-conn/make_stcon.c:1209-1455; stcon_remove_pred, stcon_pred_local_host,
 * st-conn/make_stcon_func.cu:238-392). stats as fc_stcon, time including the
 * path assembly on the device. */
fc_status fc_stcon_path(fc_graph* g, int32_t s, int32_t t, int32_t* path_out, int32_t cap,
                        int32_t* len_out, fc_stats* stats);

/* Helper-seed s–t connectivity (BFS_4K_MULTI,
 * BFS_4K_MULTI/src/device/BFS_WorkEfficient.cu:9-168 with the contact
 * closure of .../WorkEfficientKernel/BFS_WE_AdjMatrixAnalysis.cu:1-35):
 * s, t and `helpers` random seeds (0..510, distinct, from `seed`) each grow
 * a colour in one shared frontier; contacts between colours are closed
 * from s's colour after every level. connected: 1 when the closure reaches
 * t's colour, 0 when the frontier empties first. A verdict only: the
 * chain of helper balls is no shortest path. stats: time_ms, levels,
 * visited_vertices. */
fc_status fc_stcon_seeds(fc_graph* g, int32_t s, int32_t t, int32_t helpers, uint64_t seed,
                         int32_t* connected, fc_stats* stats);

/* ST-CON's check of a matching node mn (validate_stcon_bfs,
 * st-conn/make_stcon.c:1578-1641): a BFS rooted at mn; following its tree
 * parents from t and from s must reach mn within depth steps each, and the
 * longer walk must take exactly depth steps (depth = ST-CON's DEPTH,
 * ceil(dist/2)). FC_OK when it holds, FC_ERR_MISMATCH with the reason
 * otherwise. */
fc_status fc_validate_stcon(fc_graph* g, int32_t s, int32_t t, int32_t mn, int32_t depth);

/* Scheduled one-query-per-launch visits. Only FC_POLICY_PLAIN is built (it
 * is fc_stcon_many); FC_POLICY_HEURISTIC and FC_POLICY_LEARNED return
 * FC_ERR_UNSUPPORTED. Heuristic and learned scheduling are available in the
 * batched engine, fc_stcon_batch. */
fc_status fc_stcon_sched_many(fc_graph* g, const int32_t* s, const int32_t* t, int32_t n, int32_t policy,
                              int32_t* connected, int32_t* dist, fc_stats* stats);

/* Scheduling policy of a visit step. In fc_stcon_batch, the heuristic and
 * learned policies also run bridge steps: before a side whose live frontier
 * has at least 1 M edges (FASTCON_BRIDGE_MIN_EDGES) expands, the edges between
 * the two sides' current frontiers are tested directly, which answers the
 * queries they connect (distance = sum of the two levels + 1). Graphs whose
 * device lists are not id-sorted (FC_GRAPH_KEEP_ORDER, FC_GRAPH_KEEP_IDS)
 * run without them. Distances are the same under every policy. */
typedef enum fc_policy {
  FC_POLICY_PLAIN = 0,      /* both colours, top-down, every iteration (reference semantics) */
  FC_POLICY_HEURISTIC = 1,  /* smaller side first; direction and width from frontier statistics */
  FC_POLICY_LEARNED = 2     /* the trained MLP picks side, direction and width; heuristic fallback */
} fc_policy;

typedef struct fc_batch_opts {
  int32_t policy;        /* fc_policy */
  int32_t vw;            /* lanes per vertex for top-down expansion (4..64); 0 = autotuned */
  int32_t tile;          /* hub chunk: edges of the hub lists one wave takes at a time, a
                            multiple of 64 in [64, 65536]; 0 = autotuned */
  int32_t use_graph;     /* 1: replay captured hipGraphs of 16 phase launches; 0: plain launches */
  int32_t words;         /* 64-bit query words per vertex: 1 (64 queries per wave) or 0 (the
                            same); other values are FC_ERR_UNSUPPORTED */
  float   learn_margin;  /* learned policy: the network overrides the heuristic's action only
                            when it predicts a log-cost lower by at least this; < 0 = 1.0, the
                            margin chosen on validation graphs for the shipped parameters (at
                            which the network follows the heuristic on held-out graphs) */
  int32_t lanes;         /* waves run concurrently, each on its own stream and workspace
                            (1..4); 0 = default (3) */
  int32_t reserved;
} fc_batch_opts;

typedef struct fc_batch_stats {
  double  time_ms;
  int32_t waves;
  int32_t levels;              /* sum over waves of visit steps */
  int64_t active_vertices;     /* sum over steps of vertices with a non-zero frontier word */
  int64_t active_edges;        /* sum over steps of their degrees */
  int64_t touched_vertices;    /* sum over waves of vertices whose masks were written */
  int32_t vw;                  /* width used */
  int32_t tile;                /* tile schedule used */
  int32_t filters;             /* frontier-list filter passes (resolved queries dropped) */
  int32_t phases;              /* kernel launches that did work */
  int64_t launches;            /* kernel launches issued (phases + no-op tail launches) */
  int32_t bridges;             /* bridge steps: edges between the two sides' frontiers tested
                                  directly (heuristic and learned schedules) */
  int32_t bridged;             /* queries a bridge step answered */
} fc_batch_stats;

/* Batched s–t: queries are run as bit-parallel waves of 64 (one uint64 mask
 * per vertex and side). Verdicts and distances equal fc_stcon's. */
fc_status fc_stcon_batch(fc_graph* g, const int32_t* s, const int32_t* t, int32_t n,
                         int32_t* connected, int32_t* dist, const fc_batch_opts* opts,
                         fc_batch_stats* stats);

/* ------------------------------------------------------------------------ */
/* Learned scheduler: 12 frontier features -> 32 -> 32 -> 9 choices, FP32.   */
/* ------------------------------------------------------------------------ */

#define FC_SCHED_FEATURES 12
#define FC_SCHED_HIDDEN 32
#define FC_SCHED_ACTIONS 9
#define FC_SCHED_PARAMS (12 * 32 + 32 + 32 * 32 + 32 + 32 * FC_SCHED_ACTIONS + FC_SCHED_ACTIONS)

/* Scheduler actions of a batched step: sides {s, t, both}[a % 3], and by
 * a / 3 the direction and the virtual-warp width of top-down steps:
 * 0 top-down with the engine's width (the autotuned one, else chosen from
 * the frontier's mean degree), 1 top-down with a whole wave per frontier
 * vertex (width 64), 2 bottom-up. */
#define FC_ACTION_S_TOPDOWN 0
#define FC_ACTION_T_TOPDOWN 1
#define FC_ACTION_BOTH_TOPDOWN 2
#define FC_ACTION_S_TOPDOWN_WIDE 3
#define FC_ACTION_T_TOPDOWN_WIDE 4
#define FC_ACTION_BOTH_TOPDOWN_WIDE 5
#define FC_ACTION_S_BOTTOMUP 6
#define FC_ACTION_T_BOTTOMUP 7
#define FC_ACTION_BOTH_BOTTOMUP 8

/* Per-step trace record logged by timing every alternative from one state:
 * cost_us[a] is the device time from taking action a at this step to the
 * end of the wave, the rest of the wave following the traced schedule. */
typedef struct fc_trace_row {
  float features[FC_SCHED_FEATURES];
  float cost_us[FC_SCHED_ACTIONS];   /* microseconds */
  int32_t heuristic;                 /* the heuristic's action at this step */
  int32_t wave;                      /* wave index within the traced batch */
} fc_trace_row;

/* Run the batched engine along the heuristic (or learned, policy
 * FC_POLICY_LEARNED, no fallback) schedule, pausing at every
 * `every`-th step decision to time all FC_SCHED_ACTIONS alternatives (device
 * state is copied and restored around each), until cap rows are written;
 * the rest runs untraced. Returns the number written in *n_rows; dist_out
 * (nullable) receives the distances. */
fc_status fc_sched_trace(fc_graph* g, const int32_t* s, const int32_t* t, int32_t n, int32_t policy, int32_t every,
                         fc_trace_row* rows, int32_t cap, int32_t* n_rows, int32_t* dist_out);
/* Row weighting of the training loss: every row alike, or each row by the
 * time at stake, its smallest cost_us over the actions divided by the mean
 * of that over the rows (so that the loss follows total, not per-step,
 * regret). */
#define FC_SCHED_WEIGHT_UNIFORM 0
#define FC_SCHED_WEIGHT_TIME 1
/* Train on the device (hand-written forward/backward/Adam kernels), full
 * batch, from a seeded initialisation. Targets are log(cost_us) minus the
 * row's mean over the actions; the loss is
 * 0.5*sum_r w_r*sum_a (y-target)^2 / sum_r w_r with the weighting's w_r. */
fc_status fc_sched_train(fc_graph* g, const fc_trace_row* rows, int32_t n_rows, int32_t epochs,
                         float lr, uint64_t seed, int32_t weighting, float* final_loss);
fc_status fc_sched_set_params(fc_graph* g, const float* params, int32_t n);
fc_status fc_sched_get_params(fc_graph* g, float* params, int32_t n);
/* One forward pass, or the loss 0.5*sum_r w_r*sum_a (y-target)^2 / sum_r w_r
 * and its gradient w.r.t. params (weight nullable: every w_r = 1), on the
 * device; used to check the kernels against a CPU oracle. */
fc_status fc_sched_forward(fc_graph* g, const float* x, int32_t n, float* y_out);
fc_status fc_sched_gradient(fc_graph* g, const float* x, const float* target, const float* weight, int32_t n,
                            float* loss_out, float* grad_out);

#ifdef __cplusplus
}
#endif

#endif /* FASTCON_H */
