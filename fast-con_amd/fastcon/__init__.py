"""Python binding of libfastcon (ctypes over the C ABI in include/fastcon.h).

This is the host-side mirror used by the tests and the benchmark. It loads
the in-tree ``fast-con_amd/libfastcon.so`` and raises ``ImportError`` when the
library has not been built: there is no fallback implementation.

Reference interfaces mirrored (paths relative to the reference checkout):
  Graph.bfs            cudaGraph::Reset + cudaBFS4K (BFS-4K/src/device/BFS_WorkEfficient.cu:81-133)
  Graph.stcon          make_stcon (bernaschi/ST_CONN_ONE_ATOMIC/current-st/st-conn/make_stcon.c:730-1497)
  read_graph           readGraph::readGraphHeader + readSTD (BFS-4K/Util/readGraph.h:17-23)
  csr_from_edges       run_make_struct (st-conn/make_struct.c:304-356,553-738)
"""
from __future__ import annotations

import atexit
import ctypes as C
import json
import os
import weakref
from dataclasses import dataclass
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libfastcon.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() (make -C fast-con_amd)")
_lib = C.CDLL(LIB_PATH)

FC_INF = -1
(OK, ERR_INVALID, ERR_HIP, ERR_NOMEM, ERR_IO, ERR_FORMAT, ERR_CAPACITY, ERR_NOGPU, ERR_TIMEOUT, ERR_UNSUPPORTED,
 ERR_MISMATCH) = range(11)
POLICY_PLAIN, POLICY_HEURISTIC, POLICY_LEARNED = 0, 1, 2
GRAPH_KEEP_ORDER = 1
GRAPH_KEEP_IDS = 2
DIR_FILE, DIR_DIRECTED, DIR_UNDIRECTED = 0, 1, 2
SCHED_FEATURES, SCHED_HIDDEN, SCHED_ACTIONS = 12, 32, 9
SCHED_PARAMS = 12 * 32 + 32 + 32 * 32 + 32 + 32 * SCHED_ACTIONS + SCHED_ACTIONS
SCHED_WEIGHT_UNIFORM, SCHED_WEIGHT_TIME = 0, 1

_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_f32p = C.POINTER(C.c_float)


class FastconError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"fastcon status {status}: {msg}")
        self.status = status


class _HostGraph(C.Structure):
    _fields_ = [("V", C.c_int64), ("E", C.c_int64), ("offsets", _i64p), ("columns", _i32p),
                ("directed", C.c_int32), ("reserved", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("time_ms", C.c_double), ("levels", C.c_int32), ("depth", C.c_int32),
                ("visited_vertices", C.c_int64), ("visited_edges", C.c_int64),
                ("expanded_vertices", C.c_int64), ("expanded_edges", C.c_int64),
                ("topdown_levels", C.c_int32), ("bottomup_levels", C.c_int32),
                ("max_frontier", C.c_int32), ("launches", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchOpts(C.Structure):
    _fields_ = [("policy", C.c_int32), ("vw", C.c_int32), ("tile", C.c_int32), ("use_graph", C.c_int32),
                ("words", C.c_int32), ("learn_margin", C.c_float), ("lanes", C.c_int32), ("reserved", C.c_int32)]


class BatchStats(C.Structure):
    _fields_ = [("time_ms", C.c_double), ("waves", C.c_int32), ("levels", C.c_int32),
                ("active_vertices", C.c_int64), ("active_edges", C.c_int64), ("touched_vertices", C.c_int64),
                ("vw", C.c_int32), ("tile", C.c_int32), ("filters", C.c_int32), ("phases", C.c_int32),
                ("launches", C.c_int64), ("bridges", C.c_int32), ("bridged", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TraceRow(C.Structure):
    _fields_ = [("features", C.c_float * SCHED_FEATURES), ("cost_us", C.c_float * SCHED_ACTIONS),
                ("heuristic", C.c_int32), ("wave", C.c_int32)]


def _sig(name, res, *args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_lib.fc_last_error.restype = C.c_char_p
_lib.fc_version.restype = C.c_char_p
_sig("fc_host_graph_free", None, C.POINTER(_HostGraph))
_sig("fc_free", None, C.c_void_p)
_sig("fc_read_graph", C.c_int, C.c_char_p, C.c_int, C.POINTER(_HostGraph))
_sig("fc_read_edge_list", C.c_int, C.c_char_p, C.POINTER(_i64p), _i64p, _i64p)
_sig("fc_csr_from_edges", C.c_int, _i64p, C.c_int64, C.c_int64, C.POINTER(_HostGraph))
_sig("fc_gen_rmat", C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_uint64,
     C.POINTER(_HostGraph))
_sig("fc_gen_grid", C.c_int, C.c_int32, C.c_int32, C.c_double, C.c_uint64, C.POINTER(_HostGraph))
_sig("fc_gen_er", C.c_int, C.c_int64, C.c_double, C.c_uint64, C.POINTER(_HostGraph))
_sig("fc_gen_rmat_gpu", C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_uint64,
     C.POINTER(_HostGraph))
_sig("fc_csr_from_edges_gpu", C.c_int, _i64p, C.c_int64, C.c_int64, C.POINTER(_HostGraph))
_sig("fc_sample_roots", C.c_int, C.POINTER(_HostGraph), C.c_int32, C.c_uint64, _i32p)
_sig("fc_components", C.c_int, C.POINTER(_HostGraph), _i32p)
_sig("fc_graph_create_csr32", C.c_int, _i32p, _i32p, C.c_int32, C.c_int64, C.c_uint32, C.POINTER(C.c_void_p))
_sig("fc_graph_create_csr64", C.c_int, _i64p, _i32p, C.c_int64, C.c_int64, C.c_uint32, C.POINTER(C.c_void_p))
_sig("fc_graph_destroy", C.c_int, C.c_void_p)
_sig("fc_bfs", C.c_int, C.c_void_p, C.c_int32, _i32p, _i32p, C.POINTER(Stats))
_sig("fc_graph_levels", C.c_int, C.c_void_p, _i32p)
_sig("fc_graph_parents", C.c_int, C.c_void_p, _i32p)
_sig("fc_stcon", C.c_int, C.c_void_p, C.c_int32, C.c_int32, _i32p, _i32p, _i32p, C.POINTER(Stats))
_sig("fc_stcon_many", C.c_int, C.c_void_p, _i32p, _i32p, C.c_int32, _i32p, _i32p, C.POINTER(Stats))
_sig("fc_stcon_many_static", C.c_int, C.c_void_p, C.c_void_p, C.c_int32, C.c_int64, _i32p, _i32p, C.c_int32, _i32p,
     _i32p, C.POINTER(Stats))
_sig("fc_stcon_path", C.c_int, C.c_void_p, C.c_int32, C.c_int32, _i32p, C.c_int32, _i32p, C.POINTER(Stats))
_sig("fc_stcon_seeds", C.c_int, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_uint64, _i32p, C.POINTER(Stats))
_sig("fc_validate_stcon", C.c_int, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32)
_sig("fc_validate_bfs", C.c_int, C.POINTER(_HostGraph), C.c_int32, _i32p, C.c_int64, _i32p)
_sig("fc_stcon_sched_many", C.c_int, C.c_void_p, _i32p, _i32p, C.c_int32, C.c_int32, _i32p, _i32p,
     C.POINTER(Stats))
_sig("fc_stcon_batch", C.c_int, C.c_void_p, _i32p, _i32p, C.c_int32, _i32p, _i32p, C.POINTER(BatchOpts),
     C.POINTER(BatchStats))
_sig("fc_sched_trace", C.c_int, C.c_void_p, _i32p, _i32p, C.c_int32, C.c_int32, C.c_int32, C.POINTER(TraceRow),
     C.c_int32, _i32p, _i32p)
_sig("fc_sched_train", C.c_int, C.c_void_p, C.POINTER(TraceRow), C.c_int32, C.c_int32, C.c_float, C.c_uint64,
     C.c_int32, _f32p)
_sig("fc_sched_set_params", C.c_int, C.c_void_p, _f32p, C.c_int32)
_sig("fc_sched_get_params", C.c_int, C.c_void_p, _f32p, C.c_int32)
_sig("fc_sched_forward", C.c_int, C.c_void_p, _f32p, C.c_int32, _f32p)
_sig("fc_sched_gradient", C.c_int, C.c_void_p, _f32p, _f32p, _f32p, C.c_int32, _f32p, _f32p)

# Every symbol the C header declares (checked by the CPU tests).
EXPORTED = [
    "fc_last_error", "fc_version", "fc_host_graph_free", "fc_read_graph", "fc_read_edge_list", "fc_free",
    "fc_csr_from_edges", "fc_gen_rmat", "fc_gen_grid", "fc_gen_er", "fc_sample_roots", "fc_components",
    "fc_graph_create_csr32", "fc_graph_create_csr64", "fc_graph_destroy", "fc_bfs", "fc_graph_levels",
    "fc_graph_parents", "fc_stcon", "fc_stcon_many", "fc_stcon_sched_many", "fc_stcon_batch",
    "fc_sched_trace", "fc_sched_train", "fc_sched_set_params", "fc_sched_get_params", "fc_sched_forward",
    "fc_sched_gradient", "fc_stcon_path", "fc_validate_stcon", "fc_validate_bfs", "fc_gen_rmat_gpu",
    "fc_csr_from_edges_gpu", "fc_stcon_seeds", "fc_stcon_many_static",
]


def lib():
    return _lib


def version() -> str:
    return _lib.fc_version().decode()


def _check(st: int):
    if st != OK:
        raise FastconError(st, _lib.fc_last_error().decode(errors="replace"))


def _p32(a):
    return a.ctypes.data_as(_i32p)


def _p64(a):
    return a.ctypes.data_as(_i64p)


class HostGraph:
    """A CSR graph in host memory (int64 offsets, int32 columns)."""

    def __init__(self, offsets: np.ndarray, columns: np.ndarray, directed: bool = False):
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        self.columns = np.ascontiguousarray(columns, dtype=np.int32)
        self.directed = bool(directed)

    @property
    def V(self) -> int:
        return len(self.offsets) - 1

    @property
    def E(self) -> int:
        return int(self.offsets[-1])

    def degree(self) -> np.ndarray:
        return np.diff(self.offsets)

    @classmethod
    def _take(cls, hg: _HostGraph) -> "HostGraph":
        V, E = hg.V, hg.E
        off = np.ctypeslib.as_array(hg.offsets, shape=(V + 1,)).copy()
        col = np.ctypeslib.as_array(hg.columns, shape=(max(E, 1),))[:E].copy()
        d = bool(hg.directed)
        _lib.fc_host_graph_free(C.byref(hg))
        return cls(off, col, d)

    def _struct(self) -> _HostGraph:
        return _HostGraph(self.V, self.E, _p64(self.offsets), _p32(self.columns), int(self.directed), 0)

    def sample_roots(self, n: int, seed: int) -> np.ndarray:
        out = np.empty(n, np.int32)
        hg = self._struct()
        _check(_lib.fc_sample_roots(C.byref(hg), n, seed, _p32(out)))
        return out

    def components(self) -> np.ndarray:
        out = np.empty(self.V, np.int32)
        hg = self._struct()
        _check(_lib.fc_components(C.byref(hg), _p32(out)))
        return out

    def validate_bfs(self, root: int, parents, nvisited: int = -1):
        """Graph500 BFS-tree check (fc_validate_bfs). Returns (ok, message,
        depths); depths is None unless ok."""
        parents = np.ascontiguousarray(parents, np.int32)
        if len(parents) != self.V:
            raise ValueError("parents must have V entries")
        depth = np.empty(self.V, np.int32)
        hg = self._struct()
        st = _lib.fc_validate_bfs(C.byref(hg), int(root), _p32(parents), int(nvisited), _p32(depth))
        if st == OK:
            return True, "", depth
        if st == ERR_MISMATCH:
            return False, _lib.fc_last_error().decode(errors="replace"), None
        _check(st)


def gen_rmat(scale: int, edgefactor: float = 16.0, a: float = 0.57, b: float = 0.19, c: float = 0.19,
             seed: int = 1) -> HostGraph:
    hg = _HostGraph()
    _check(_lib.fc_gen_rmat(scale, edgefactor, a, b, c, seed, C.byref(hg)))
    return HostGraph._take(hg)


def gen_grid(rows: int, cols: int, p_delete: float = 0.0, seed: int = 1) -> HostGraph:
    hg = _HostGraph()
    _check(_lib.fc_gen_grid(rows, cols, p_delete, seed, C.byref(hg)))
    return HostGraph._take(hg)


def gen_er(n: int, mean_degree: float, seed: int = 1) -> HostGraph:
    hg = _HostGraph()
    _check(_lib.fc_gen_er(n, mean_degree, seed, C.byref(hg)))
    return HostGraph._take(hg)


def csr_from_edges(edges: np.ndarray, V: int = 0) -> HostGraph:
    e = np.ascontiguousarray(edges, dtype=np.int64).reshape(-1)
    hg = _HostGraph()
    _check(_lib.fc_csr_from_edges(_p64(e), len(e) // 2, V, C.byref(hg)))
    return HostGraph._take(hg)


def gen_rmat_gpu(scale: int, edgefactor: float = 16.0, a: float = 0.57, b: float = 0.19, c: float = 0.19,
                 seed: int = 1) -> HostGraph:
    """gen_rmat with generation and CSR build on the GPU (same graph)."""
    hg = _HostGraph()
    _check(_lib.fc_gen_rmat_gpu(scale, edgefactor, a, b, c, seed, C.byref(hg)))
    return HostGraph._take(hg)


def csr_from_edges_gpu(edges: np.ndarray, V: int = 0) -> HostGraph:
    """csr_from_edges with the CSR built on the GPU (same arrays)."""
    e = np.ascontiguousarray(edges, dtype=np.int64).reshape(-1)
    hg = _HostGraph()
    _check(_lib.fc_csr_from_edges_gpu(_p64(e), len(e) // 2, V, C.byref(hg)))
    return HostGraph._take(hg)


def stcon_many_static(d_off: int, d_col: int, V: int, E: int, s, t) -> StResult:
    """fc_stcon_many_static: plain s–t queries on a CSR graph in the
    caller's device memory (d_off, d_col: device addresses of V+1 and E
    int32 values), with no device allocation by the call."""
    s = np.ascontiguousarray(s, np.int32)
    t = np.ascontiguousarray(t, np.int32)
    conn = np.empty(len(s), np.int32)
    dist = np.empty(len(s), np.int32)
    st = Stats()
    _check(_lib.fc_stcon_many_static(C.c_void_p(d_off), C.c_void_p(d_col), V, E, _p32(s), _p32(t), len(s),
                                     _p32(conn), _p32(dist), C.byref(st)))
    return StResult(conn, dist, st.as_dict())


def read_graph(path: str, direction: int = DIR_FILE) -> HostGraph:
    hg = _HostGraph()
    _check(_lib.fc_read_graph(path.encode(), direction, C.byref(hg)))
    return HostGraph._take(hg)


def read_edge_list(path: str):
    p = _i64p()
    m = C.c_int64()
    nv = C.c_int64()
    _check(_lib.fc_read_edge_list(path.encode(), C.byref(p), C.byref(m), C.byref(nv)))
    arr = np.ctypeslib.as_array(p, shape=(max(2 * m.value, 1),))[: 2 * m.value].copy()
    _lib.fc_free(C.cast(p, C.c_void_p))
    return arr.reshape(-1, 2), nv.value


# Device graphs still open. They are destroyed when the interpreter exits,
# from an atexit hook: Python runs those before the C runtime's exit
# handlers, so the HIP runtime is still up when their streams, graphs and
# buffers are released. A graph left to the garbage collector at shutdown
# could otherwise be released after the runtime has been torn down.
_live_graphs = weakref.WeakSet()


@atexit.register
def _close_live_graphs():
    for g in list(_live_graphs):
        g.close()


@dataclass
class StResult:
    connected: np.ndarray
    dist: np.ndarray
    stats: dict


class Graph:
    """A graph resident on the GPU (library-owned device copy + workspace)."""

    def __init__(self, hg: HostGraph, force64: bool = False, keep_order: bool = False, keep_ids: bool = False):
        """keep_order: keep the caller's adjacency order on the device
        (default: neighbours by decreasing degree); keep_ids: keep the
        caller's vertex numbering on the device (default: renumbered by
        decreasing degree). Calls take and return caller ids either way."""
        self.V, self.E = hg.V, hg.E
        self._h = C.c_void_p()
        flags = (GRAPH_KEEP_ORDER if keep_order else 0) | (GRAPH_KEEP_IDS if keep_ids else 0)
        if hg.E < 2**31 - 1 and not force64:
            off32 = hg.offsets.astype(np.int32)
            _check(_lib.fc_graph_create_csr32(_p32(off32), _p32(hg.columns), hg.V, hg.E, flags, C.byref(self._h)))
            self.offsets64 = False
        else:
            _check(_lib.fc_graph_create_csr64(_p64(hg.offsets), _p32(hg.columns), hg.V, hg.E, flags,
                                              C.byref(self._h)))
            self.offsets64 = True
        _live_graphs.add(self)

    def close(self):
        if self._h:
            _lib.fc_graph_destroy(self._h)
            self._h = C.c_void_p()
        _live_graphs.discard(self)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def bfs(self, src: int, levels: bool = True, parents: bool = False):
        lv = np.empty(self.V, np.int32) if levels else None
        pa = np.empty(self.V, np.int32) if parents else None
        st = Stats()
        _check(_lib.fc_bfs(self._h, src, _p32(lv) if lv is not None else None,
                           _p32(pa) if pa is not None else None, C.byref(st)))
        return lv, pa, st.as_dict()

    def levels(self) -> np.ndarray:
        lv = np.empty(self.V, np.int32)
        _check(_lib.fc_graph_levels(self._h, _p32(lv)))
        return lv

    def parents(self) -> np.ndarray:
        pa = np.empty(self.V, np.int32)
        _check(_lib.fc_graph_parents(self._h, _p32(pa)))
        return pa

    def stcon(self, s: int, t: int):
        conn, dist, meet = C.c_int32(), C.c_int32(), C.c_int32()
        st = Stats()
        _check(_lib.fc_stcon(self._h, s, t, C.byref(conn), C.byref(dist), C.byref(meet), C.byref(st)))
        return bool(conn.value), dist.value, meet.value, st.as_dict()

    def stcon_path(self, s: int, t: int, cap: int = 4096):
        """A shortest s–t path (fc_stcon_path): (path as int32 array, s
        first; empty when not connected), stats."""
        for _ in range(2):
            out = np.empty(max(cap, 1), np.int32)
            n = C.c_int32()
            st = Stats()
            code = _lib.fc_stcon_path(self._h, int(s), int(t), _p32(out), cap, C.byref(n), C.byref(st))
            if code == ERR_CAPACITY and n.value > cap:
                cap = n.value
                continue
            _check(code)
            return out[: n.value].copy(), st.as_dict()
        _check(code)

    def stcon_seeds(self, s: int, t: int, helpers: int = 318, seed: int = 1):
        """Helper-seed verdict (fc_stcon_seeds): (connected, stats)."""
        conn = C.c_int32()
        st = Stats()
        _check(_lib.fc_stcon_seeds(self._h, int(s), int(t), int(helpers), int(seed), C.byref(conn), C.byref(st)))
        return bool(conn.value), st.as_dict()

    def validate_stcon(self, s: int, t: int, mn: int, depth: int):
        """ST-CON's matching-node check (fc_validate_stcon): (ok, message)."""
        code = _lib.fc_validate_stcon(self._h, int(s), int(t), int(mn), int(depth))
        if code == OK:
            return True, ""
        if code == ERR_MISMATCH:
            return False, _lib.fc_last_error().decode(errors="replace")
        _check(code)

    def stcon_many(self, s, t, policy: int = POLICY_PLAIN) -> StResult:
        s = np.ascontiguousarray(s, np.int32)
        t = np.ascontiguousarray(t, np.int32)
        conn = np.empty(len(s), np.int32)
        dist = np.empty(len(s), np.int32)
        st = Stats()
        if policy == POLICY_PLAIN:
            _check(_lib.fc_stcon_many(self._h, _p32(s), _p32(t), len(s), _p32(conn), _p32(dist), C.byref(st)))
        else:
            _check(_lib.fc_stcon_sched_many(self._h, _p32(s), _p32(t), len(s), policy, _p32(conn), _p32(dist),
                                            C.byref(st)))
        return StResult(conn, dist, st.as_dict())

    # margin of the learned policy used when a call gives none (negative:
    # the library's default); load_default_sched_params sets the shipped one
    learn_margin = -1.0

    def stcon_batch(self, s, t, policy: int = POLICY_HEURISTIC, vw: int = 0, tile: int = 0, use_graph: int = 1,
                    words: int = 1, learn_margin: float = None, lanes: int = 0) -> StResult:
        if learn_margin is None:
            learn_margin = self.learn_margin
        s = np.ascontiguousarray(s, np.int32)
        t = np.ascontiguousarray(t, np.int32)
        conn = np.empty(len(s), np.int32)
        dist = np.empty(len(s), np.int32)
        o = BatchOpts(policy, vw, tile, use_graph, words, learn_margin, lanes)
        st = BatchStats()
        _check(_lib.fc_stcon_batch(self._h, _p32(s), _p32(t), len(s), _p32(conn), _p32(dist), C.byref(o),
                                   C.byref(st)))
        return StResult(conn, dist, st.as_dict())

    # ---- learned scheduler ----
    def sched_trace(self, s, t, cap: int = 4096, every: int = 1, policy: int = POLICY_HEURISTIC):
        """Trace the heuristic schedule over the queries. Returns (features
        [rows, 12], cost_us [rows, 9], heuristic action [rows], wave [rows],
        distances [n])."""
        s = np.ascontiguousarray(s, np.int32)
        t = np.ascontiguousarray(t, np.int32)
        rows = (TraceRow * max(cap, 1))()
        n = C.c_int32()
        dist = np.empty(len(s), np.int32)
        _check(_lib.fc_sched_trace(self._h, _p32(s), _p32(t), len(s), policy, every, rows, cap, C.byref(n),
                                   _p32(dist)))
        k = n.value
        arr = np.ctypeslib.as_array(rows)[:k] if k else None
        feats = np.zeros((k, SCHED_FEATURES), np.float32)
        cost = np.zeros((k, SCHED_ACTIONS), np.float32)
        heur = np.zeros(k, np.int32)
        wave = np.zeros(k, np.int32)
        for i in range(k):
            feats[i] = rows[i].features[:]
            cost[i] = rows[i].cost_us[:]
            heur[i] = rows[i].heuristic
            wave[i] = rows[i].wave
        return feats, cost, heur, wave, dist

    @staticmethod
    def _rows(feats: np.ndarray, cost: np.ndarray):
        rows = (TraceRow * len(feats))()
        for i in range(len(feats)):
            rows[i].features[:] = [float(x) for x in feats[i]]
            rows[i].cost_us[:] = [float(x) for x in cost[i]]
        return rows

    def sched_train(self, feats: np.ndarray, cost: np.ndarray, epochs: int = 2000, lr: float = 3e-3,
                    seed: int = 1, weighting: int = SCHED_WEIGHT_UNIFORM) -> float:
        """Train the scheduler network on the device from traced rows; the
        trained parameters are set on this graph. weighting is
        SCHED_WEIGHT_UNIFORM or SCHED_WEIGHT_TIME (rows weighted by their
        smallest cost over the mean of those). Returns the final loss."""
        rows = self._rows(feats, cost)
        loss = C.c_float()
        _check(_lib.fc_sched_train(self._h, rows, len(feats), epochs, lr, seed, weighting, C.byref(loss)))
        return loss.value

    def sched_set_params(self, w: np.ndarray):
        w = np.ascontiguousarray(w, np.float32)
        _check(_lib.fc_sched_set_params(self._h, w.ctypes.data_as(_f32p), len(w)))

    def sched_get_params(self) -> np.ndarray:
        w = np.empty(SCHED_PARAMS, np.float32)
        _check(_lib.fc_sched_get_params(self._h, w.ctypes.data_as(_f32p), len(w)))
        return w

    def sched_forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        y = np.empty((len(x), SCHED_ACTIONS), np.float32)
        _check(_lib.fc_sched_forward(self._h, x.ctypes.data_as(_f32p), len(x), y.ctypes.data_as(_f32p)))
        return y

    def sched_gradient(self, x: np.ndarray, target: np.ndarray, weight: Optional[np.ndarray] = None):
        """Loss 0.5*sum_r w_r*sum_a (y-target)^2 / sum_r w_r and its gradient
        (weight None: every w_r = 1)."""
        x = np.ascontiguousarray(x, np.float32)
        tg = np.ascontiguousarray(target, np.float32)
        wp = None
        if weight is not None:
            wt = np.ascontiguousarray(weight, np.float32)
            if wt.shape != (len(x),):
                raise ValueError("weight must have one entry per row")
            wp = wt.ctypes.data_as(_f32p)
        g = np.empty(SCHED_PARAMS, np.float32)
        loss = C.c_float()
        _check(_lib.fc_sched_gradient(self._h, x.ctypes.data_as(_f32p), tg.ctypes.data_as(_f32p), wp, len(x),
                                      C.byref(loss), g.ctypes.data_as(_f32p)))
        return loss.value, g


# Trained scheduler parameters shipped with the package (written by
# fastcon/train_sched.py); absent until a training run has been committed.
DEFAULT_SCHED_PARAMS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sched_params.npy")


def load_default_sched_params(g: "Graph", path: str = DEFAULT_SCHED_PARAMS) -> bool:
    """Set the shipped scheduler parameters on g, and the margin chosen with
    them (sched_report.json beside them) as g's default; False when there
    are none."""
    if not os.path.exists(path):
        return False
    w = np.load(path, allow_pickle=False)
    if w.dtype != np.float32 or w.shape != (SCHED_PARAMS,):
        raise ValueError(f"{path}: expected {SCHED_PARAMS} float32 parameters")
    g.sched_set_params(w)
    rep = os.path.join(os.path.dirname(path), "sched_report.json")
    if os.path.exists(rep):
        with open(rep) as fh:
            m = json.load(fh).get("margin")
        if isinstance(m, (int, float)) and m >= 0:
            g.learn_margin = float(m)
    return True
