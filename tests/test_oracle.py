"""Pins the CPU oracle: closed forms, brute force, planted mistakes and the
reference's published s16 known-answer file. No GPU."""
import json
import os

import numpy as np
import pytest

from conftest import small_graphs

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def floyd(off, col):
    V = len(off) - 1
    INF = 10**9
    D = np.full((V, V), INF, dtype=np.int64)
    np.fill_diagonal(D, 0)
    for v in range(V):
        for e in range(off[v], off[v + 1]):
            D[v, col[e]] = min(D[v, col[e]], 1)
    for k in range(V):
        D = np.minimum(D, D[:, k : k + 1] + D[k : k + 1, :])
    D[D >= INF] = -1
    return D


def test_bfs_closed_forms(fc, orc):
    # path: level = |i - src|
    g = fc.csr_from_edges(np.array([[i, i + 1] for i in range(29)]), 30)
    d, q, info = orc.bfs(g.offsets, g.columns, 7)
    assert list(d) == [abs(i - 7) for i in range(30)]
    assert info == (30, g.E, 22)
    # grid without deletions: Manhattan distance
    R, Cc = 11, 17
    g = fc.gen_grid(R, Cc, 0.0, 3)
    src = 5 * Cc + 9
    d, _, _ = orc.bfs(g.offsets, g.columns, src)
    exp = [abs(v // Cc - 5) + abs(v % Cc - 9) for v in range(R * Cc)]
    assert list(d) == exp
    # star: centre 0, leaves at 1, leaf-to-leaf 2
    g = fc.csr_from_edges(np.array([[0, i] for i in range(1, 9)]), 9)
    d, _, _ = orc.bfs(g.offsets, g.columns, 3)
    assert list(d) == [1, 2, 2, 0, 2, 2, 2, 2, 2]


def test_bfs_unreached_encoding_and_counts(fc, orc):
    g = fc.csr_from_edges(np.array([[0, 1], [1, 2], [3, 4]]), 6)
    d, q, info = orc.bfs(g.offsets, g.columns, 0, inf=orc.REF_INF)
    assert list(d) == [0, 1, 2, orc.REF_INF, orc.REF_INF, orc.REF_INF]
    # visitedEdges = sum of degrees of dequeued vertices when not all reached
    assert info == (3, 4, 2)
    assert list(q) == [0, 1, 2]


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_bfs_vs_floyd(fc, orc, seed):
    rng = np.random.default_rng(seed)
    V = 40
    m = int(rng.integers(20, 80))
    e = rng.integers(0, V, size=(m, 2))
    g = fc.csr_from_edges(e, V)
    D = floyd(g.offsets, g.columns)
    for s in range(0, V, 7):
        d, _, _ = orc.bfs(g.offsets, g.columns, s)
        assert np.array_equal(d, D[s].astype(np.int32))
        for t in range(V):
            c, dist, depth = orc.stcon(g.offsets, g.columns, s, t)
            assert dist == D[s, t] and c == (D[s, t] >= 0)
            assert depth == ((D[s, t] + 1) // 2 if D[s, t] >= 0 else -1)


def test_to_csr_order_matches_reference_rules(orc):
    # Graph::ToCSR places entries in COO order; undirected adds the reverse right after.
    coo = np.array([[2, 0], [0, 1], [1, 2], [0, 2], [0, 0]])
    off, col = orc.to_csr(3, coo, undirected=True)
    assert list(off) == [0, 5, 7, 10]
    assert list(col) == [2, 1, 2, 0, 0, 0, 2, 0, 1, 0]
    off, col = orc.to_csr(3, coo, undirected=False)
    assert list(off) == [0, 3, 4, 5]
    assert list(col) == [1, 2, 0, 2, 0]


def test_canonical_is_set_semantics(orc):
    rng = np.random.default_rng(9)
    V = 50
    e = rng.integers(-1, V, size=(400, 2))
    off, col = orc.canonical(e, V)
    model = set()
    for u, v in e:
        if u < 0 or v < 0 or u == v:
            continue
        model.add((u, v))
        model.add((v, u))
    got = {(v, int(col[k])) for v in range(V) for k in range(off[v], off[v + 1])}
    assert got == model and len(col) == len(model)
    for v in range(V):
        assert np.all(np.diff(col[off[v] : off[v + 1]]) > 0)


def test_validate_bfs_tree_rejects_planted_mistakes(fc, orc):
    g = fc.gen_grid(6, 6, 0.0, 1)
    lv, _, _ = orc.bfs(g.offsets, g.columns, 0)
    par = np.full(36, -1, np.int32)
    par[0] = 0
    for v in range(1, 36):
        for k in range(g.offsets[v], g.offsets[v + 1]):
            u = g.columns[k]
            if lv[u] == lv[v] - 1:
                par[v] = u
                break
    assert orc.validate_bfs_tree(g.offsets, g.columns, 0, lv, par) == 0
    bad = par.copy()
    bad[14] = 35  # not a neighbour, wrong level
    assert orc.validate_bfs_tree(g.offsets, g.columns, 0, lv, bad) == 2
    bad = lv.copy()
    bad[20] += 1
    assert orc.validate_bfs_tree(g.offsets, g.columns, 0, bad, par) != 0
    bad = par.copy()
    bad[0] = 1
    assert orc.validate_bfs_tree(g.offsets, g.columns, 0, lv, bad) == 1


def test_count_visited_edges_and_statistics(fc, orc):
    g = fc.gen_grid(5, 8, 0.0, 1)
    lv, _, _ = orc.bfs(g.offsets, g.columns, 0)
    assert orc.count_visited_edges(g.offsets, g.columns, lv) == g.E // 2
    x = np.array([3.0, 1.0, 4.0, 1.0, 5.0, 9.0, 2.0, 6.0])
    r = orc.get_statistics(x)
    assert r["min"] == 1.0 and r["max"] == 9.0
    assert abs(r["mean"] - x.mean()) < 1e-12 and abs(r["std"] - x.std(ddof=1)) < 1e-12
    xs = np.sort(x)
    assert r["median"] == (xs[3] + xs[4]) / 2


def test_stcon_work_brute_force(fc, orc):
    rng = np.random.default_rng(3)
    g = fc.gen_er(60, 2.5, 4)
    D = floyd(g.offsets, g.columns)
    deg = np.diff(g.offsets)
    for _ in range(20):
        s, t = rng.integers(0, 60, 2)
        if s == t:
            continue
        vx, ex = orc.stcon_work(g.offsets, g.columns, s, t)
        d = D[s, t]
        if d >= 0:
            Dh = (d + 1) // 2
            sel = [v for v in range(60) if (0 <= D[s, v] <= Dh - 1) or (0 <= D[t, v] <= Dh - 1)]
        else:
            sel = [v for v in range(60) if D[s, v] >= 0 or D[t, v] >= 0]
        assert vx == len(sel) and ex == int(deg[sel].sum())


def test_batch_wave_distances(fc, orc):
    for name, g in small_graphs(fc):
        rng = np.random.default_rng(len(name))
        V = g.V
        s = rng.integers(0, V, 40).astype(np.int32)
        t = rng.integers(0, V, 40).astype(np.int32)
        dist, A, Es, k = orc.batch_wave(g.offsets, g.columns, s, t)
        for i in range(40):
            assert dist[i] == orc.stcon(g.offsets, g.columns, s[i], t[i])[1], name


def _load_known_answer():
    return json.load(open(os.path.join(GOLD, "st_stats_s16.json")))


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                                    "oracle", "_ref", "libgraph500.so")),
                    reason="reference Graph500 generator not built in oracle/_ref")
def test_known_answer_s16(orc):
    """The reference's published ST-CON run (st_stats_20200418_140644:9-108):
    its 100 (RED-V, BLUE-V) pairs and DEPTH = ceil(d/2) from the host BFS."""
    fx = _load_known_answer()
    V = 1 << fx["scale"]
    e = orc.graph500_edges(fx["scale"], fx["edgefactor"], fx["seed1"], fx["seed2"])
    off, col = orc.canonical(e, V)
    assert len(col) == 1831864
    roots = orc.find_stcon_roots(2 * len(fx["rows"]), off, col, *fx["root_seeds"])
    n = len(fx["rows"])
    maxd = 0
    for i, (run, red, blue, depth, mn) in enumerate(fx["rows"]):
        assert (roots[i], roots[i + n]) == (red, blue)
        c, d, dp = orc.stcon(off, col, red, blue)
        assert c == (mn > 0) and dp == depth
        maxd = max(maxd, dp)
    assert maxd == fx["max_level"]


def test_count_visited_edges_hand_case(orc):
    """SC/adj_func.c:334-357 counts, for each reached vertex v, its
    adjacency entries u <= v. Hand-worked: undirected edges (0,1), (1,2), two
    self-loops (2,2) and (0,0) kept by Graph::ToCSR, and a second component
    {3,4,5} the BFS from 0 never reaches. Reached lists: 0 -> [1,0,0],
    1 -> [0,2], 2 -> [1,2,2], so 2 + 1 + 3 = 6 (each non-loop edge once,
    each self-loop's two entries twice). Writing < for <= gives 2, and
    counting unreached vertices too gives 8."""
    coo = np.array([[0, 1], [1, 2], [2, 2], [0, 0], [3, 4], [4, 5]], np.int32)
    off, col = orc.to_csr(6, coo, True)
    lv, _, _ = orc.bfs(off, col, 0)
    assert list(lv) == [0, 1, 2, -1, -1, -1]
    assert orc.count_visited_edges(off, col, lv) == 6
    lv3, _, _ = orc.bfs(off, col, 3)
    assert orc.count_visited_edges(off, col, lv3) == 2  # edges (3,4) and (4,5) once each


def test_batch_wave_closed_form_counters(fc, orc):
    """The bit-parallel wave on graphs whose frontiers are known in closed
    form. Path 0..n-1, one query (0, n-1): iteration k expands vertex k-1 on
    the s side and n-k on the t side, so with m = ceil((n-1)/2) iterations
    A = 2m frontier (vertex, side) pairs and E = 4m - 2 edges (both ends
    have degree 1, inner vertices 2), and the distance is n-1. Star with
    centre 0 and 20 leaves plus an isolated vertex 21: leaf-to-leaf queries
    meet at the centre in one iteration (distance 2, even), centre-to-leaf
    queries in one (distance 1, odd); a query to the isolated vertex is not
    connected after one iteration, the t side having gained nothing."""
    for n in range(2, 13):
        e = np.array([[i, i + 1] for i in range(n - 1)])
        g = fc.csr_from_edges(e, n)
        dist, A, E, k = orc.batch_wave(g.offsets, g.columns, np.array([0], np.int32), np.array([n - 1], np.int32))
        m = n // 2
        assert (dist[0], A, E, k) == (n - 1, 2 * m, 4 * m - 2, m), n
    e = np.array([[0, i] for i in range(1, 21)])
    g = fc.csr_from_edges(e, 22)
    s = np.array([1, 2, 3, 0, 4], np.int32)
    t = np.array([5, 6, 7, 9, 21], np.int32)
    dist, A, E, k = orc.batch_wave(g.offsets, g.columns, s, t)
    assert list(dist) == [2, 2, 2, 1, -1] and k == 1
    # frontier (vertex, side) pairs: s side {1,2,3,0,4}, t side {5,6,7,9,21}
    assert A == 10
    # degrees: leaves 1, centre 20, isolated 0
    assert E == 4 * 1 + 20 + 4 * 1 + 0


def test_validate_bfs_tree_rules_3_and_4(fc, orc):
    """Rule 4: an unreached vertex must have no parent. Rule 3: an edge may
    not join a reached and an unreached vertex, nor two levels more than one
    apart."""
    e = np.array([[0, 1], [1, 2], [2, 3], [5, 6]])
    g = fc.csr_from_edges(e, 7)
    lv, _, _ = orc.bfs(g.offsets, g.columns, 0)
    par = np.array([0, 0, 1, 2, -1, -1, -1], np.int32)
    assert orc.validate_bfs_tree(g.offsets, g.columns, 0, lv, par) == 0
    bad = par.copy()
    bad[5] = 6  # an unreached vertex with a parent
    assert orc.validate_bfs_tree(g.offsets, g.columns, 0, lv, bad) == 4
    lv2, par2 = lv.copy(), par.copy()
    lv2[3], par2[3] = -1, -1  # 3 dropped from the tree while its neighbour 2 is in it
    assert orc.validate_bfs_tree(g.offsets, g.columns, 0, lv2, par2) == 3
    e4 = np.array([[0, 1], [1, 2], [2, 3], [0, 3]])
    g4 = fc.csr_from_edges(e4, 4)
    lv4, _, _ = orc.bfs(g4.offsets, g4.columns, 0)
    assert list(lv4) == [0, 1, 2, 1]
    par4 = np.array([0, 0, 1, 0], np.int32)
    assert orc.validate_bfs_tree(g4.offsets, g4.columns, 0, lv4, par4) == 0
    # vertex 3 claimed at level 3 through parent 2: the tree edge is fine but
    # the edge (0,3) now joins levels 0 and 3
    lv4b = lv4.copy()
    lv4b[3] = 3
    par4b = np.array([0, 0, 1, 2], np.int32)
    assert orc.validate_bfs_tree(g4.offsets, g4.columns, 0, lv4b, par4b) == 3
