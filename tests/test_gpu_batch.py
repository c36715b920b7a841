"""GPU parity of the batched bit-parallel s–t engine: verdicts and distances
equal the host BFS (oracle) and the plain one-query-per-launch path, under
every schedule, width, tile and launch mode; the plain schedule's work
counters equal the oracle's bit-parallel model."""
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


def _pairs(hg, n, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, hg.V, n).astype(np.int32), rng.integers(0, hg.V, n).astype(np.int32)


@pytest.mark.parametrize("policy", [0, 1])
def test_batch_small_graphs(fc, orc, policy):
    for name, hg in small_graphs(fc):
        s, t = _pairs(hg, 100, hg.V + 1)
        s[:3] = t[:3]  # s == t cases
        exp = _oracle_dist(orc, hg, s, t)
        with fc.Graph(hg) as g:
            say(f"batch {name} policy={policy}")
            res = g.stcon_batch(s, t, policy=policy, vw=0, tile=256)
            bad = np.flatnonzero(res.dist != exp)
            assert bad.size == 0, f"{name}: {bad[:5]} got {res.dist[bad[:5]]} want {exp[bad[:5]]}"
            assert np.array_equal(res.connected, (exp >= 0).astype(np.int32))


def test_batch_rmat16_matches_plain_and_oracle(fc, orc):
    hg = fc.gen_rmat(16, 16, seed=16)
    r = hg.sample_roots(512, 17)
    s, t = r[:256], r[256:]
    exp = _oracle_dist(orc, hg, s, t)
    with fc.Graph(hg) as g:
        plain = g.stcon_many(s, t)
        assert np.array_equal(plain.dist, exp)
        for policy in (0, 1):
            for vw, tile, ug in [(0, 0, 1), (4, 256, 0), (8, 1024, 1), (32, 256, 1), (64, 512, 0)]:
                say(f"batch rmat16 policy={policy} vw={vw} tile={tile} graph={ug}")
                res = g.stcon_batch(s, t, policy=policy, vw=vw, tile=tile, use_graph=ug)
                assert np.array_equal(res.dist, exp), (policy, vw, tile, ug)


def test_batch_disconnected_and_grid(fc, orc):
    hg = fc.gen_er(5000, 1.2, 21)
    s, t = _pairs(hg, 130, 5)
    exp = _oracle_dist(orc, hg, s, t)
    assert (exp < 0).sum() > 10
    with fc.Graph(hg) as g:
        for policy in (0, 1):
            res = g.stcon_batch(s, t, policy=policy, vw=8, tile=256)
            assert np.array_equal(res.dist, exp), policy
    hg = fc.gen_grid(120, 90, 0.05, 3)
    s, t = _pairs(hg, 64, 9)
    exp = _oracle_dist(orc, hg, s, t)
    with fc.Graph(hg) as g:
        for policy in (0, 1):
            res = g.stcon_batch(s, t, policy=policy, vw=4, tile=256)
            assert np.array_equal(res.dist, exp), policy


def test_batch_plain_schedule_work_matches_oracle_model(fc, orc):
    """The plain schedule advances both sides every step; its frontier-vertex
    and frontier-edge counters equal the oracle's bit-parallel restatement."""
    hg = fc.gen_rmat(13, 8, seed=4)
    r = hg.sample_roots(128, 2)
    s, t = r[:64], r[64:]
    dist, A, Es, k = orc.batch_wave(hg.offsets, hg.columns, s, t)
    with fc.Graph(hg) as g:
        res = g.stcon_batch(s, t, policy=0, vw=16, tile=256, use_graph=0)
        assert np.array_equal(res.dist, dist)
        assert res.stats["active_vertices"] == A
        assert res.stats["active_edges"] == Es
        assert res.stats["levels"] == k


def test_batch_reuse_and_errors(fc, orc):
    hg = fc.gen_rmat(11, 8, seed=7)
    r = hg.sample_roots(200, 3)
    with fc.Graph(hg) as g:
        exp = _oracle_dist(orc, hg, r[:100], r[100:])
        for _ in range(3):
            assert np.array_equal(g.stcon_batch(r[:100], r[100:], vw=8, tile=256).dist, exp)
        # plain queries on the same graph after batches, and batches after them
        assert np.array_equal(g.stcon_many(r[:100], r[100:]).dist, exp)
        assert np.array_equal(g.stcon_batch(r[:100], r[100:], vw=0, tile=0).dist, exp)
        with pytest.raises(fc.FastconError) as ei:
            g.stcon_batch([0], [hg.V])
        assert ei.value.status == fc.ERR_INVALID
        with pytest.raises(fc.FastconError):
            g.stcon_batch([0], [1], vw=7)
        assert len(g.stcon_batch([], []).dist) == 0


def _forced_params(fc, action):
    """Scheduler parameters whose network always picks `action` (all weights
    zero, lowest output bias on that action)."""
    w = np.zeros(fc.SCHED_PARAMS, np.float32)
    w[-fc.SCHED_ACTIONS + action] = -1.0
    return w


@pytest.mark.parametrize("action", range(9))
def test_batch_every_action_is_exact(fc, orc, action):
    """Each scheduler action (side s / t / both; top-down at the engine's
    width, top-down at width 64, bottom-up) forced
    at every step still gives the oracle's verdicts and distances."""
    cases = [("rmat13", fc.gen_rmat(13, 8, seed=31)), ("er-sparse", fc.gen_er(3000, 1.3, 8)),
             ("grid", fc.gen_grid(60, 50, 0.04, 2))]
    for name, hg in cases:
        s, t = _pairs(hg, 150, action + 3)
        s[:2] = t[:2]
        exp = _oracle_dist(orc, hg, s, t)
        with fc.Graph(hg) as g:
            g.sched_set_params(_forced_params(fc, action))
            say(f"batch forced action={action} {name}")
            res = g.stcon_batch(s, t, policy=fc.POLICY_LEARNED, vw=0, tile=256)
            bad = np.flatnonzero(res.dist != exp)
            assert bad.size == 0, f"{name}: {bad[:5]} got {res.dist[bad[:5]]} want {exp[bad[:5]]}"


def test_batch_lanes_agree(fc, orc):
    """Waves split over 1..4 concurrent lanes (streams and workspaces) give
    the same answers, including batches with fewer waves than lanes."""
    hg = fc.gen_rmat(14, 16, seed=41)
    r = hg.sample_roots(900, 42)
    s, t = r[:450], r[450:]
    exp = _oracle_dist(orc, hg, s, t)
    with fc.Graph(hg) as g:
        for lanes in (1, 2, 3, 4):
            for n in (450, 100, 64, 1):
                say(f"batch lanes={lanes} n={n}")
                res = g.stcon_batch(s[:n], t[:n], lanes=lanes)
                assert np.array_equal(res.dist, exp[:n]), (lanes, n)
        with pytest.raises(fc.FastconError):
            g.stcon_batch(s, t, lanes=5)


def test_batch_launch_surplus_stays_bounded(fc, orc):
    """Launches are issued in captured graphs of 16 against a running
    estimate of the working launches per wave; over many calls the surplus
    (no-op launches after the last phase) stays within one graph and a
    margin per lane instead of growing call by call."""
    hg = fc.gen_rmat(14, 16, seed=43)
    r = hg.sample_roots(1024, 44)
    s, t = r[:512], r[512:]
    exp = _oracle_dist(orc, hg, s, t)
    with fc.Graph(hg) as g:
        surplus = []
        for k in range(60):
            if k % 20 == 0:
                say(f"batch launch surplus call {k}")
            res = g.stcon_batch(s, t, lanes=3)
            assert np.array_equal(res.dist, exp)
            surplus.append(res.stats["launches"] - res.stats["phases"])
        lanes = 3
        assert max(surplus[-20:]) <= lanes * 24, surplus
        assert np.mean(surplus[-20:]) <= np.mean(surplus[5:25]) + 16, surplus
