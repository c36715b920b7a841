"""Helper-seed s–t verdicts (fc_stcon_seeds, BFS_4K_MULTI's method) against
the host BFS, on graphs with many components (sparse ER, grids with many
deleted edges, R-MAT with isolated and small components), with 0, 62 and
318 helper seeds and both vertex numberings."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _pairs(hg, n, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, hg.V, n).astype(np.int32), rng.integers(0, hg.V, n).astype(np.int32)


@pytest.mark.parametrize("helpers", [0, 62, 318])
@pytest.mark.parametrize("keep_ids", [False, True])
def test_seed_verdicts_match_host_bfs(fc, orc, helpers, keep_ids):
    graphs = [("er", fc.gen_er(20000, 1.3, 5)), ("grid", fc.gen_grid(96, 96, 0.4, 6)),
              ("rmat", fc.gen_rmat(13, 4, seed=7))]
    for name, hg in graphs:
        s, t = _pairs(hg, 24, 11)
        s[0] = t[0]  # s == t
        comp = hg.components()
        exp = comp[s] == comp[t]
        with fc.Graph(hg, keep_ids=keep_ids) as g:
            got = []
            for a, b in zip(s, t):
                c, st = g.stcon_seeds(int(a), int(b), helpers=helpers, seed=3)
                got.append(c)
            got = np.array(got)
        say(f"{name} helpers={helpers} keep_ids={keep_ids}: {exp.sum()} of {len(s)} connected")
        assert np.array_equal(got, exp), name
        assert 0 < exp.sum() < len(s), "the case needs both verdicts"


def test_seed_levels_shrink_on_a_grid(fc, orc):
    """On a high-diameter grid the helper balls meet after far fewer levels
    than the plain two-colour visit needs."""
    hg = fc.gen_grid(256, 256, 0.0, 1)
    s, t = 0, hg.V - 1  # opposite corners, distance 510
    with fc.Graph(hg) as g:
        c0, d, _, st0 = g.stcon(s, t)
        c1, st1 = g.stcon_seeds(s, t, helpers=318, seed=2)
    assert c0 and c1 and d == 510
    say(f"grid256 corners: plain {st0['levels']} levels {st0['time_ms']:.2f} ms; "
        f"seeds {st1['levels']} levels {st1['time_ms']:.2f} ms")
    assert st1["levels"] * 4 < st0["levels"]


def test_seed_tiny_graphs(fc, orc):
    """More helpers than vertices, isolated vertices and an edgeless graph."""
    cases = [
        (fc.csr_from_edges(np.array([[0, 1]]), 2), [(0, 1, True)]),
        (fc.csr_from_edges(np.array([[0, 1], [2, 3]]), 5), [(0, 1, True), (1, 2, False), (4, 0, False), (3, 2, True)]),
        (fc.csr_from_edges(np.zeros((0, 2), np.int64), 4), [(0, 3, False), (2, 2, True)]),
    ]
    for hg, qs in cases:
        with fc.Graph(hg) as g:
            for s, t, exp in qs:
                for helpers in (0, 5, 510):
                    c, _ = g.stcon_seeds(s, t, helpers=helpers, seed=1)
                    assert c == exp, (hg.V, s, t, helpers)
