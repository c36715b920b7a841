"""GPU parity of the batched engine's bridge step.

Before a side expands, a bridge step tests the edges between the two
sides' current frontiers directly and answers the queries it connects at
distance i + j + 1 (i, j the sides' levels). It runs under the heuristic
and learned schedules, not under the plain one.

* Small graphs with the edge threshold at 1, so that every eligible step
  is preceded by a bridge: distances equal the host BFS (oracle) for every
  query, and bridges happen.
* R-MAT s20 at the default threshold: the heuristic schedule, with
  bridges, gives the plain schedule's distances and the host BFS's for the
  queries of eight sources."""
import os
import sys

import numpy as np
import pytest

from conftest import small_graphs

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _oracle_dist(orc, hg, s, t):
    return np.array([orc.stcon(hg.offsets, hg.columns, a, b)[1] for a, b in zip(s, t)], np.int32)


class _Env:
    def __init__(self, **kv):
        self.kv = kv

    def __enter__(self):
        self.old = {k: os.environ.get(k) for k in self.kv}
        os.environ.update({k: str(v) for k, v in self.kv.items()})

    def __exit__(self, *a):
        for k, v in self.old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


@pytest.mark.parametrize("lanes", [1, 3])
def test_bridge_small_graphs_match_oracle(fc, orc, lanes):
    total = 0
    with _Env(FASTCON_BRIDGE_MIN_EDGES=1):
        for name, hg in small_graphs(fc):
            rng = np.random.default_rng(hg.V + 7)
            s = rng.integers(0, hg.V, 192).astype(np.int32)
            t = rng.integers(0, hg.V, 192).astype(np.int32)
            exp = _oracle_dist(orc, hg, s, t)
            with fc.Graph(hg) as g:
                for pol in (fc.POLICY_HEURISTIC, fc.POLICY_LEARNED):
                    if pol == fc.POLICY_LEARNED and not fc.load_default_sched_params(g):
                        continue
                    say(f"bridge {name} policy={pol} lanes={lanes}")
                    res = g.stcon_batch(s, t, policy=pol, lanes=lanes)
                    bad = np.flatnonzero(res.dist != exp)
                    assert bad.size == 0, f"{name}: {bad[:5]} got {res.dist[bad[:5]]} want {exp[bad[:5]]}"
                    assert res.stats["bridged"] <= len(s)
                    total += res.stats["bridges"]
    assert total > 0, "no bridge step ran"


@pytest.mark.timeout(300)
def test_bridge_rmat20_default_threshold(fc, orc):
    hg = fc.gen_rmat(20, 16, 0.57, 0.19, 0.19, seed=2020)
    r = hg.sample_roots(2048, 2021)
    s, t = r[:1024].copy(), r[1024:].copy()
    src = r[:8]
    s[:256] = np.repeat(src, 32)
    exp = np.empty(256, np.int32)
    for k, root in enumerate(src):
        d, _, _ = orc.bfs(hg.offsets, hg.columns, int(root))
        exp[32 * k:32 * (k + 1)] = d[t[32 * k:32 * (k + 1)]]
    with fc.Graph(hg) as g:
        h = g.stcon_batch(s, t, policy=fc.POLICY_HEURISTIC)
        p = g.stcon_batch(s, t, policy=fc.POLICY_PLAIN)
        assert np.array_equal(h.dist, p.dist), "heuristic (with bridges) and plain schedules differ"
        assert np.array_equal(h.dist[:256], exp), "distances differ from the host BFS"
        assert p.stats["bridges"] == 0
        assert h.stats["bridges"] > 0 and h.stats["bridged"] > 0
        say(f"rmat20 1024 queries: {h.stats['bridges']} bridges answered {h.stats['bridged']} queries; "
            f"heuristic {h.stats['time_ms']:.2f} ms, plain {p.stats['time_ms']:.2f} ms")
