"""Host-side engine logic (no GPU): generators, canonicaliser, readers,
samplers, and the exported C ABI. Checked against the oracle."""
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def test_library_exports_every_declared_symbol(fc):
    hdr = open(os.path.join(ROOT, "include", "fastcon.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    declared = set(re.findall(r"\b(fc_[a-z0-9_]+)\s*\(", hdr))
    assert declared, "no declarations parsed"
    lib = fc.lib()
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    assert declared <= set(fc.EXPORTED), sorted(declared - set(fc.EXPORTED))
    assert fc.version().startswith("fastcon")


def test_csr_from_edges_equals_oracle(fc, orc):
    rng = np.random.default_rng(1)
    for V, m in [(1, 0), (5, 3), (100, 1000), (2000, 30000)]:
        e = rng.integers(-1, V, size=(m, 2)) if m else np.zeros((0, 2), np.int64)
        g = fc.csr_from_edges(e, V)
        off, col = orc.canonical(e, V)
        assert np.array_equal(g.offsets, off) and np.array_equal(g.columns, col)


def test_csr_from_edges_rejects_out_of_range(fc):
    with pytest.raises(fc.FastconError) as ei:
        fc.csr_from_edges(np.array([[0, 5]]), 5)
    assert ei.value.status == fc.ERR_INVALID


def _symmetric_sorted_no_loops(g):
    off, col = g.offsets, g.columns
    V = g.V
    src = np.repeat(np.arange(V), np.diff(off))
    assert not np.any(src == col), "self-loop"
    for v in range(0, V, max(1, V // 500)):
        a = col[off[v] : off[v + 1]]
        assert np.all(np.diff(a) > 0), "unsorted or duplicate"
    fw = set(zip(src.tolist(), col.tolist()))
    assert all((b, a) in fw for a, b in list(fw)[:20000])


def test_rmat_generator(fc):
    g = fc.gen_rmat(12, 16, seed=3)
    assert g.V == 4096
    _symmetric_sorted_no_loops(g)
    g2 = fc.gen_rmat(12, 16, seed=3)
    assert np.array_equal(g.offsets, g2.offsets) and np.array_equal(g.columns, g2.columns)
    g3 = fc.gen_rmat(12, 16, seed=4)
    assert not np.array_equal(g.columns[:1000], g3.columns[:1000])
    # skew: R-MAT with a=0.57 has a heavy tail
    deg = g.degree()
    assert deg.max() > 20 * deg.mean()
    # the generated multigraph has 16*4096 pairs; after symmetrisation the
    # directed count is at most twice that
    assert g.E <= 2 * 16 * 4096 and g.E > 16 * 4096


def test_rmat_quadrant_rules(fc):
    """Degenerate initiators pin the quadrant -> bit mapping: a=1 makes every
    pair (0,0) (a self-loop, dropped); b=1 makes every pair (0, N-1), one
    undirected edge; d=1 gives (N-1, N-1), dropped; a=b=0.5 fixes the row at 0
    and draws the column uniformly, a star around the image of vertex 0."""
    N = 1 << 10
    assert fc.gen_rmat(10, 4, 1.0, 0.0, 0.0, seed=1).E == 0
    g = fc.gen_rmat(10, 4, 0.0, 1.0, 0.0, seed=1)
    assert g.E == 2
    assert fc.gen_rmat(10, 4, 0.0, 0.0, 0.0, seed=1).E == 0
    g = fc.gen_rmat(10, 4, 0.5, 0.5, 0.0, seed=1)
    deg = g.degree()
    hub = int(np.argmax(deg))
    assert np.count_nonzero(deg) == deg[hub] + 1
    expect = N * (1 - np.exp(-4.0))  # distinct columns among 4N uniform draws
    assert abs(deg[hub] - expect) < 0.05 * N
    # c=1: every pair (N-1, 0), the same single edge as b=1 up to direction
    assert fc.gen_rmat(10, 4, 0.0, 0.0, 1.0, seed=1).E == 2


def test_grid_generator(fc):
    g = fc.gen_grid(30, 20, 0.0, 1)
    assert g.V == 600 and g.E == 2 * (30 * 19 + 29 * 20)
    gd = fc.gen_grid(200, 200, 0.02, 5)
    und = gd.E // 2
    full = 200 * 199 * 2
    assert abs((full - und) / full - 0.02) < 0.004
    _symmetric_sorted_no_loops(gd)
    e = []
    for v in range(gd.V):
        for k in range(gd.offsets[v], gd.offsets[v + 1]):
            u = gd.columns[k]
            assert abs(u - v) in (1, 200)


def test_er_generator(fc):
    g = fc.gen_er(10000, 8.0, 2)
    _symmetric_sorted_no_loops(g)
    assert abs(g.E / 10000 - 8.0) < 0.2


def test_sample_roots(fc):
    g = fc.gen_rmat(10, 4, seed=2)
    r = g.sample_roots(200, 7)
    assert len(set(r.tolist())) == 200
    assert np.all(g.degree()[r] > 0)
    assert np.array_equal(r, g.sample_roots(200, 7))
    with pytest.raises(fc.FastconError):
        fc.csr_from_edges(np.array([[0, 1]]), 3).sample_roots(3, 1)


def test_components(fc, orc):
    g = fc.gen_er(2000, 1.0, 3)
    lab = g.components()
    for s in range(0, 2000, 97):
        d, _, _ = orc.bfs(g.offsets, g.columns, s)
        assert np.array_equal(lab == lab[s], d >= 0)
        assert lab[s] == np.flatnonzero(d >= 0).min()


# ---------------------------------------------------------------- readers
def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_read_matrix_market(fc, orc, tmp_path):
    txt = "%%MatrixMarket matrix coordinate pattern symmetric\n% c\n4 4 4\n1 2\n2 3 7.5\n4 1\n3 3\n"
    g = fc.read_graph(_write(tmp_path, "a.mtx", txt))
    off, col = orc.to_csr(4, np.array([[0, 1], [1, 2], [3, 0], [2, 2]]), True)
    assert np.array_equal(g.offsets, off) and np.array_equal(g.columns, col) and not g.directed
    g = fc.read_graph(_write(tmp_path, "b.mtx", txt.replace("symmetric", "general")))
    off, col = orc.to_csr(4, np.array([[0, 1], [1, 2], [3, 0], [2, 2]]), False)
    assert np.array_equal(g.columns, col) and g.directed
    g = fc.read_graph(_write(tmp_path, "c.mtx", txt.replace("symmetric", "general")), fc.DIR_UNDIRECTED)
    assert g.E == 8


def test_read_snap_remaps_first_seen(fc, orc, tmp_path):
    txt = "# Undirected graph\n# Nodes: 4 Edges: 3\n# FromNodeId ToNodeId\n100 7\n7 55\n9 100\n"
    g = fc.read_graph(_write(tmp_path, "s.txt", txt))
    # ids by first appearance: 100->0, 7->1, 55->2, 9->3
    off, col = orc.to_csr(4, np.array([[0, 1], [1, 2], [3, 0]]), True)
    assert np.array_equal(g.offsets, off) and np.array_equal(g.columns, col)
    g = fc.read_graph(_write(tmp_path, "d.txt", txt.replace("Undirected", "Directed")))
    assert g.directed and g.E == 3


def test_read_dimacs9_and_dimacs10(fc, orc, tmp_path):
    txt = "c gtgraph\np sp 3 3\na 1 2 5\nc mid\na 2 3 1\na 3 1 1\n"
    g = fc.read_graph(_write(tmp_path, "g.gr", txt))
    off, col = orc.to_csr(3, np.array([[0, 1], [1, 2], [2, 0]]), False)
    assert np.array_equal(g.columns, col) and g.directed
    # METIS / DIMACS10: one line of 1-based neighbours per vertex, kept as listed
    txt = "% metis\n4 3\n2 3\n1\n1 4\n3\n"
    g = fc.read_graph(_write(tmp_path, "m.graph", txt))
    assert list(g.offsets) == [0, 2, 3, 5, 6] and list(g.columns) == [1, 2, 0, 0, 3, 2]
    assert not g.directed


def test_read_errors(fc, tmp_path):
    with pytest.raises(fc.FastconError) as ei:
        fc.read_graph(str(tmp_path / "missing.mtx"))
    assert ei.value.status == fc.ERR_IO
    with pytest.raises(fc.FastconError) as ei:
        fc.read_graph(str(tmp_path))
    assert ei.value.status == fc.ERR_IO
    with pytest.raises(fc.FastconError) as ei:
        fc.read_graph(_write(tmp_path, "x.txt", "hello world\n1 2\n"))
    assert ei.value.status == fc.ERR_FORMAT
    with pytest.raises(fc.FastconError) as ei:
        fc.read_graph(_write(tmp_path, "p.txt", "## pg\n2 2 1\n0 1\n"))
    assert ei.value.status == fc.ERR_FORMAT


def test_read_edge_list(fc, tmp_path):
    txt = "# Directed graph\n# Nodes: 6 Edges: 3\n# FromNodeId\tToNodeId\n0 5\n3 2\n5 5\n"
    e, nv = fc.read_edge_list(_write(tmp_path, "e.txt", txt))
    assert nv == 6 and e.tolist() == [[0, 5], [3, 2], [5, 5]]


def _tree_case(fc, orc, hg, root):
    """Oracle levels and a BFS tree built from them (each reached vertex's
    smallest neighbour one level up)."""
    lv, _, _ = orc.bfs(hg.offsets, hg.columns, root)
    pa = np.full(hg.V, -1, np.int32)
    pa[root] = root
    off, col = hg.offsets, hg.columns
    for v in np.flatnonzero(lv > 0):
        nb = col[off[v]:off[v + 1]]
        pa[v] = nb[lv[nb] == lv[v] - 1].min()
    return lv, pa


def test_validate_bfs_accepts_oracle_trees_and_names_planted_mistakes(fc, orc):
    """fc_validate_bfs (Graph500 validate_bfs_result) accepts the oracle's
    BFS trees, returns their depths, and rejects each planted mistake with
    the reference's message."""
    hg = fc.gen_rmat(12, 8, seed=3)
    roots = hg.sample_roots(3, 4)
    for r in roots:
        lv, pa = _tree_case(fc, orc, hg, int(r))
        ok, msg, depth = hg.validate_bfs(int(r), pa, nvisited=int((pa != -1).sum()))
        assert ok, msg
        assert np.array_equal(depth, lv)
        assert orc.validate_bfs_tree(hg.offsets, hg.columns, int(r), lv, pa) == 0
    r = int(roots[0])
    lv, pa = _tree_case(fc, orc, hg, r)
    reached = np.flatnonzero((lv >= 2))
    v = int(reached[0])
    cases = []
    bad = pa.copy(); bad[r] = int(pa[v]); cases.append(("root", bad, "parent of root vertex"))
    bad = pa.copy(); bad[v] = v; cases.append(("self", bad, "is itself"))
    bad = pa.copy(); bad[v] = hg.V + 5; cases.append(("range", bad, "out-of-range"))
    # a parent that is not a neighbour
    nb = set(hg.columns[hg.offsets[v]:hg.offsets[v + 1]].tolist())
    far = next(int(x) for x in np.flatnonzero(lv >= 0) if int(x) not in nb and int(x) != v)
    bad = pa.copy(); bad[v] = far; cases.append(("edge", bad, "no graph edge"))
    # a two-cycle between neighbours away from the root
    u = next(int(x) for x in hg.columns[hg.offsets[v]:hg.offsets[v + 1]] if int(x) != r and lv[int(x)] >= 1)
    bad = pa.copy(); bad[v] = u; bad[u] = v; cases.append(("cycle", bad, "do not form a tree"))
    # a leaf of the tree marked unreached, although its neighbours are reached
    leaf = next(int(x) for x in reached if not (pa == x).any())
    bad = pa.copy(); bad[leaf] = -1; cases.append(("outside", bad, "BFS tree"))
    for name, b, text in cases:
        ok, msg, _ = hg.validate_bfs(r, b)
        assert not ok and text in msg, (name, msg)
    ok, msg, _ = hg.validate_bfs(r, pa, nvisited=int((pa != -1).sum()) + 1)
    assert not ok and "visit count" in msg


def test_validate_bfs_depth_gap(fc, orc):
    """A parent tree that is a valid tree but not a BFS tree (an edge whose
    ends differ in depth by two) is rejected."""
    # path 0-1-2-3 plus the chord 0-2: the tree 0->1->2->3 puts 2 at depth 2
    e = np.array([[0, 1], [1, 2], [2, 3], [0, 2]])
    hg = fc.csr_from_edges(e, 4)
    ok, _, d = hg.validate_bfs(0, np.array([0, 0, 0, 2], np.int32))
    assert ok and d.tolist() == [0, 1, 1, 2]
    ok, msg, _ = hg.validate_bfs(0, np.array([0, 0, 1, 2], np.int32))
    assert not ok and "too far apart" in msg


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; --gpus above 1 exits with a message before
    touching a device."""
    import subprocess
    import sys

    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_python_constants_match_the_header(fc):
    """The ctypes mirror's constants equal the C header's #defines."""
    hdr = open(os.path.join(ROOT, "include", "fastcon.h")).read()
    defs = dict(re.findall(r"#define\s+(FC_[A-Z0-9_]+)\s+(-?\d+)u?\b", hdr))
    pairs = {"FC_SCHED_FEATURES": fc.SCHED_FEATURES, "FC_SCHED_HIDDEN": fc.SCHED_HIDDEN,
             "FC_SCHED_ACTIONS": fc.SCHED_ACTIONS, "FC_SCHED_WEIGHT_UNIFORM": fc.SCHED_WEIGHT_UNIFORM,
             "FC_SCHED_WEIGHT_TIME": fc.SCHED_WEIGHT_TIME, "FC_GRAPH_KEEP_ORDER": fc.GRAPH_KEEP_ORDER,
             "FC_GRAPH_KEEP_IDS": fc.GRAPH_KEEP_IDS}
    for name, val in pairs.items():
        assert name in defs, name
        assert int(defs[name]) == val, (name, defs[name], val)
    F, H, A = fc.SCHED_FEATURES, fc.SCHED_HIDDEN, fc.SCHED_ACTIONS
    assert fc.SCHED_PARAMS == F * H + H + H * H + H + H * A + A


def test_generators_match_golden_fixtures(fc):
    """The seeded graphs of the tests and the benchmark (R-MAT s16 and s22,
    the 4096^2 grid with 2% deletions, ER 2^20 with mean degree 8) equal the
    committed fixtures bit for bit (tests/golden/make_generator_fixtures.py),
    and their counts obey what the construction fixes: a grid keeps each of
    its 2n(n-1) undirected edges with probability 1-p (binomial, checked
    within 6 standard deviations) and an ER graph has at most n*d/2
    undirected edges."""
    import json
    import sys

    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import make_generator_fixtures as mk

    fix = json.load(open(os.path.join(ROOT, "tests", "golden", "generators.json")))
    assert [c["name"] for c in fix] == [c["name"] for c in mk.CONFIGS]
    for c in fix:
        hg = mk.build(fc, c)
        got = mk.describe(hg)
        for k, v in got.items():
            assert v == c[k], f"{c['name']}: {k} is {v}, fixture {c[k]}"
        assert sum(got["log2_degree_hist"]) == got["V"] - got["isolated"]
        if c["kind"] == "grid":
            n_und = 2 * c["rows"] * (c["cols"] - 1)
            p = c["p_delete"]
            mu, sd = n_und * (1 - p), (n_und * p * (1 - p)) ** 0.5
            assert abs(got["E"] / 2 - mu) < 6 * sd
        if c["kind"] == "er":
            assert got["E"] <= c["n"] * c["mean_degree"]
    rmat22 = [c for c in fix if c["name"] == "rmat22"][0]
    assert rmat22["E"] == 128302336  # the benchmark graph
    g = fc.gen_grid(64, 64, 0.0, 3)
    assert g.E == 4 * 64 * 63  # no deletions: every undirected edge in both directions
