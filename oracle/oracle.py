"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) and of the reference's
own Graph500 generator built in oracle/_ref/ (when the reference checkout was
present at build time).

TEST INFRASTRUCTURE ONLY. The oracle is the checker: tests/, the smoke check
and bench.py's cpu_baseline leg may use it; the engine never does.

What pins each oracle function (so that a plausible slip fails a check):
  bfs             closed forms (paths, stars, cycles, complete graphs, grids
                  without deletions -> Manhattan distance), Floyd–Warshall
                  brute force on random graphs of <= 40 vertices, the BFS
                  invariants, and the reference's published s16 known-answer
                  file (100 DEPTH values, st_stats_20200418_140644:9-108)
  to_csr          hand-worked COO examples (order of both directions)
  canonical       set semantics vs a Python set-of-pairs model; the s16
                  directed-edge count 1,831,864 of the Graph500 graph
  stcon           symmetry d(s,t) = d(t,s); the s16 known answers
  stcon_work      brute-force enumeration on small graphs; the GPU plain
                  visit's expanded counters equal it on connected pairs
  validate_bfs_tree  planted mistakes (root, wrong parent, wrong level,
                  non-edge, an unreached vertex with a parent, an edge
                  across the tree's boundary or across two levels) are each
                  rejected with their rule's code
  count_visited_edges  a hand-worked graph with self-loops and an unreached
                  component (a < for <= or a dropped reached test fails it),
                  and E/2 on a connected loop-free grid
  get_statistics  numpy mean/std(ddof=1)/min/max/quantile rules
  find_stcon_roots  the 100 (s,t) pairs of the s16 known-answer file
  batch_wave      distances equal bfs-based distances per query; its A/E and
                  iteration counters follow the closed forms of a path and
                  a star
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = C.CDLL(os.path.join(_HERE, "liboracle.so"))
_REF = os.path.join(_HERE, "_ref", "libgraph500.so")
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_f64p = C.POINTER(C.c_double)

RNG_FN = C.CFUNCTYPE(None, C.c_int64, C.c_uint64, C.c_uint64, C.c_int64, _f64p)


def _sig(name, res, *args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_sig("orc_bfs", C.c_int, C.c_int64, C.c_int64, _i64p, _i32p, C.c_int32, C.c_int32, _i32p, _i32p, _i64p)
_sig("orc_to_csr", C.c_int, C.c_int64, _i32p, C.c_int64, C.c_int, _i64p, _i32p)
_sig("orc_canonical", C.c_int64, _i64p, C.c_int64, C.c_int64, _i64p, _i32p)
_sig("orc_stcon", C.c_int, C.c_int64, C.c_int64, _i64p, _i32p, C.c_int32, C.c_int32, _i32p, _i32p, _i32p)
_sig("orc_stcon_work", C.c_int, C.c_int64, C.c_int64, _i64p, _i32p, C.c_int32, C.c_int32, _i64p, _i64p)
_sig("orc_validate_bfs_tree", C.c_int, C.c_int64, _i64p, _i32p, C.c_int32, _i32p, _i32p)
_sig("orc_count_visited_edges", C.c_int64, C.c_int64, _i64p, _i32p, _i32p)
_sig("orc_get_statistics", None, _f64p, C.c_int, _f64p)
_sig("orc_find_stcon_roots", C.c_int, C.c_int, C.c_int64, _i64p, _i32p, C.c_uint64, C.c_uint64, C.c_void_p, _i64p)
_sig("orc_batch_wave", C.c_int, C.c_int64, _i64p, _i32p, C.c_int, _i32p, _i32p, _i32p, _i64p, _i64p, _i32p)
_sig("orc_time_stcon", C.c_double, C.c_int64, C.c_int64, _i64p, _i32p, C.c_int, _i32p, _i32p, _i32p)
_sig("orc_time_bfs", C.c_double, C.c_int64, C.c_int64, _i64p, _i32p, C.c_int, _i32p, _i64p)

REF_INF = 65535  # (dist_t)-1 with dist_t = unsigned short (BFS-4K/config.h:14,50)


def _p32(a):
    return a.ctypes.data_as(_i32p)


def _p64(a):
    return a.ctypes.data_as(_i64p)


def _csr(off, col):
    return np.ascontiguousarray(off, np.int64), np.ascontiguousarray(col, np.int32)


def bfs(off, col, src, inf=-1):
    """Graph::BfsInit + Graph::bfs. Returns (dist, queue, (visitedNodes, visitedEdges, maxDistance))."""
    off, col = _csr(off, col)
    V = len(off) - 1
    dist = np.empty(V, np.int32)
    queue = np.empty(V, np.int32)
    info = np.zeros(3, np.int64)
    rc = _lib.orc_bfs(V, int(off[-1]), _p64(off), _p32(col), int(src), int(inf), _p32(dist), _p32(queue), _p64(info))
    if rc:
        raise ValueError(f"orc_bfs failed ({rc})")
    return dist, queue[: info[0]], tuple(int(x) for x in info)


def to_csr(V, coo, undirected):
    coo = np.ascontiguousarray(coo, np.int32).reshape(-1)
    n = len(coo) // 2
    E = 2 * n if undirected else n
    off = np.empty(V + 1, np.int64)
    col = np.empty(max(E, 1), np.int32)
    rc = _lib.orc_to_csr(V, _p32(coo), n, 1 if undirected else 0, _p64(off), _p32(col))
    if rc:
        raise ValueError(f"orc_to_csr failed ({rc})")
    return off, col[:E]


def canonical(edges, V):
    e = np.ascontiguousarray(edges, np.int64).reshape(-1)
    m = len(e) // 2
    off = np.empty(V + 1, np.int64)
    col = np.empty(max(2 * m, 1), np.int32)
    E = _lib.orc_canonical(_p64(e), m, V, _p64(off), _p32(col))
    if E < 0:
        raise ValueError(f"orc_canonical failed ({E})")
    return off, col[:E]


def stcon(off, col, s, t):
    off, col = _csr(off, col)
    c, d, dp = C.c_int32(), C.c_int32(), C.c_int32()
    rc = _lib.orc_stcon(len(off) - 1, int(off[-1]), _p64(off), _p32(col), int(s), int(t), C.byref(c), C.byref(d),
                        C.byref(dp))
    if rc:
        raise ValueError(f"orc_stcon failed ({rc})")
    return bool(c.value), d.value, dp.value


def stcon_work(off, col, s, t):
    off, col = _csr(off, col)
    vx, ex = C.c_int64(), C.c_int64()
    _lib.orc_stcon_work(len(off) - 1, int(off[-1]), _p64(off), _p32(col), int(s), int(t), C.byref(vx), C.byref(ex))
    return vx.value, ex.value


def validate_bfs_tree(off, col, root, level, parent):
    off, col = _csr(off, col)
    level = np.ascontiguousarray(level, np.int32)
    parent = np.ascontiguousarray(parent, np.int32)
    return _lib.orc_validate_bfs_tree(len(off) - 1, _p64(off), _p32(col), int(root), _p32(level), _p32(parent))


def count_visited_edges(off, col, level):
    off, col = _csr(off, col)
    level = np.ascontiguousarray(level, np.int32)
    return int(_lib.orc_count_visited_edges(len(off) - 1, _p64(off), _p32(col), _p32(level)))


def get_statistics(x):
    x = np.ascontiguousarray(x, np.float64)
    r = np.zeros(7)
    _lib.orc_get_statistics(x.ctypes.data_as(_f64p), len(x), r.ctypes.data_as(_f64p))
    return dict(zip(["min", "q1", "median", "q3", "max", "mean", "std"], r))


def batch_wave(off, col, s, t):
    off, col = _csr(off, col)
    s = np.ascontiguousarray(s, np.int32)
    t = np.ascontiguousarray(t, np.int32)
    assert len(s) <= 64
    dist = np.empty(len(s), np.int32)
    a, e, k = C.c_int64(), C.c_int64(), C.c_int32()
    _lib.orc_batch_wave(len(off) - 1, _p64(off), _p32(col), len(s), _p32(s), _p32(t), _p32(dist), C.byref(a),
                        C.byref(e), C.byref(k))
    return dist, a.value, e.value, k.value


def time_stcon(off, col, s, t, lib=None):
    """Single-thread host s–t (BFS from s, level of t). Returns (seconds, dist)."""
    lib = lib or _lib
    off, col = _csr(off, col)
    s = np.ascontiguousarray(s, np.int32)
    t = np.ascontiguousarray(t, np.int32)
    dist = np.empty(len(s), np.int32)
    sec = lib.orc_time_stcon(len(off) - 1, int(off[-1]), _p64(off), _p32(col), len(s), _p32(s), _p32(t), _p32(dist))
    return sec, dist


def time_bfs(off, col, roots, lib=None):
    """Single-thread host BFS from each root. Returns (seconds, sum of the
    visitedEdges counts)."""
    lib = lib or _lib
    off, col = _csr(off, col)
    roots = np.ascontiguousarray(roots, np.int32)
    tot = C.c_int64()
    sec = lib.orc_time_bfs(len(off) - 1, int(off[-1]), _p64(off), _p32(col), len(roots), _p32(roots), C.byref(tot))
    return sec, tot.value


def native_build(workdir):
    """The CPU baseline's build: oracle.c compiled with -O3 -march=native for
    the CPU this runs on, into workdir. Returns (lib, flags) for the lib=
    argument of time_stcon / time_bfs, or (None, reason) when gcc fails (the
    prebuilt -O3 liboracle.so is then the one to use)."""
    import subprocess

    flags = ["-O3", "-march=native", "-std=gnu11", "-fPIC", "-shared"]
    out = os.path.join(workdir, "liboracle_native.so")
    try:
        subprocess.run(["gcc", *flags, "-o", out, os.path.join(_HERE, "oracle.c"), "-lm"], check=True,
                       capture_output=True, timeout=120)
    except (OSError, subprocess.SubprocessError) as e:
        return None, f"gcc failed: {e}"
    lib = C.CDLL(out)
    lib.orc_time_stcon.restype = C.c_double
    lib.orc_time_stcon.argtypes = [C.c_int64, C.c_int64, _i64p, _i32p, C.c_int, _i32p, _i32p, _i32p]
    lib.orc_time_bfs.restype = C.c_double
    lib.orc_time_bfs.argtypes = [C.c_int64, C.c_int64, _i64p, _i32p, C.c_int, _i32p, _i64p]
    return lib, "gcc " + " ".join(flags[:2])


# ---------------------------------------------------------------------------
# The reference's Graph500 generator (oracle/_ref/libgraph500.so).
# ---------------------------------------------------------------------------
def have_ref() -> bool:
    return os.path.exists(_REF)


def _ref():
    lib = C.CDLL(_REF)
    lib.make_graph.restype = None
    lib.make_graph.argtypes = [C.c_int, C.c_int64, C.c_uint64, C.c_uint64, _f64p, _i64p, C.POINTER(_i64p)]
    lib.make_random_numbers.restype = None
    lib.make_random_numbers.argtypes = [C.c_int64, C.c_uint64, C.c_uint64, C.c_int64, _f64p]
    return lib


def graph500_edges(scale, edgefactor, seed1, seed2, initiator=(0.57, 0.19, 0.19, 0.05)):
    """make_graph(SCALE, edgefactor*2^SCALE, seed1, seed2, initiator) as ST-CON calls it
    (st-conn/main.c:354-374). Returns an (m, 2) int64 array; pairs of -1 are dropped edges."""
    lib = _ref()
    ini = (C.c_double * 4)(*initiator)
    n = C.c_int64()
    p = _i64p()
    lib.make_graph(scale, int(edgefactor * 2**scale), seed1, seed2, ini, C.byref(n), C.byref(p))
    arr = np.ctypeslib.as_array(p, shape=(2 * n.value,)).copy()
    C.CDLL(None).free(C.cast(p, C.c_void_p))
    return arr.reshape(-1, 2)


def find_stcon_roots(n, off, col, seed1=23, seed2=24):
    """find_stcon_roots on one rank with Graph500's make_random_numbers; seeds
    23/24 as st-conn/make_stcon.c:382."""
    off, col = _csr(off, col)
    lib = _ref()
    roots = np.empty(n, np.int64)
    fn = C.cast(lib.make_random_numbers, C.c_void_p)
    k = _lib.orc_find_stcon_roots(n, len(off) - 1, _p64(off), _p32(col), seed1, seed2, fn, _p64(roots))
    return roots[:k]
