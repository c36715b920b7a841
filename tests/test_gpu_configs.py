"""GPU parity at full scale for the two remaining workloads of the path:

* high-diameter plain s–t: 4096 x 4096 grid with 2% of its edges deleted,
  256 pairs at Manhattan distance 1,000-6,000 (thousands of levels, small
  frontiers): the plain visit and the batched engine against the host BFS;
* batched s–t on R-MAT scale 24 (int64 offsets), 16,384 queries, 70% with
  both endpoints in the giant component and 30% with one endpoint in a small
  component (uneven reach): the batched engine against the plain visit and,
  for the queries of two shared sources, the host BFS.

Distances are compared bit for bit."""
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _grid_pairs(rows, cols, nsrc, per_src, lo, hi, seed):
    rng = np.random.default_rng(seed)
    s, t = [], []
    while len(s) < nsrc * per_src:
        r0, c0 = int(rng.integers(0, rows)), int(rng.integers(0, cols))
        got = 0
        for _ in range(100000):
            if got == per_src:
                break
            dr, dc = int(rng.integers(-hi, hi + 1)), int(rng.integers(-hi, hi + 1))
            r1, c1 = r0 + dr, c0 + dc
            if lo <= abs(dr) + abs(dc) <= hi and 0 <= r1 < rows and 0 <= c1 < cols:
                s.append(r0 * cols + c0)
                t.append(r1 * cols + c1)
                got += 1
        if got < per_src:  # source too close to a corner for the range: drop it
            del s[len(s) - got:], t[len(t) - got:]
    return np.array(s, np.int32), np.array(t, np.int32)


@pytest.mark.timeout(600)
def test_grid4096_high_diameter(fc, orc):
    rows = cols = 4096
    t0 = time.time()
    hg = fc.gen_grid(rows, cols, 0.02, 7)
    s, t = _grid_pairs(rows, cols, 8, 32, 1000, 6000, 11)
    exp = np.empty(len(s), np.int32)
    for k in range(0, len(s), 32):
        d, _, _ = orc.bfs(hg.offsets, hg.columns, int(s[k]))
        exp[k:k + 32] = d[t[k:k + 32]]
    say(f"grid4096: V={hg.V} E={hg.E}, oracle for {len(s)} pairs in {time.time() - t0:.1f}s; "
        f"distances {exp.min()}..{exp.max()}, unconnected {(exp < 0).sum()}")
    with fc.Graph(hg) as g:
        plain = g.stcon_many(s, t)
        assert np.array_equal(plain.dist, exp)
        say(f"grid4096 plain: {len(s) / (plain.stats['time_ms'] / 1e3):.1f} q/s")
        for pol in (fc.POLICY_PLAIN, fc.POLICY_HEURISTIC):
            res = g.stcon_batch(s, t, policy=pol)
            assert np.array_equal(res.dist, exp), pol
            say(f"grid4096 batched policy {pol}: {len(s) / (res.stats['time_ms'] / 1e3):.1f} q/s")


@pytest.mark.timeout(600)
def test_rmat24_batched_uneven_reach(fc, orc):
    t0 = time.time()
    hg = fc.gen_rmat(24, 16, 0.57, 0.19, 0.19, seed=24)
    deg = hg.degree()
    say(f"rmat24: V={hg.V} E={hg.E} built in {time.time() - t0:.1f}s")
    with fc.Graph(hg, force64=True) as g:
        assert g.offsets64
        root = int(hg.sample_roots(1, 5)[0])
        lv, _, st = g.bfs(root)
        giant = np.flatnonzero(lv >= 0).astype(np.int32)
        small = np.flatnonzero((lv < 0) & (deg > 0)).astype(np.int32)
        assert len(giant) > hg.V // 4 and len(small) > 1000
        rng = np.random.default_rng(25)
        n = 16384
        s = rng.choice(giant, n)
        t = rng.choice(giant, n)
        far = rng.random(n) < 0.3
        t[far] = rng.choice(small, int(far.sum()))
        # the first 64 queries share two sources, so two host BFSs check them
        a, b = rng.choice(giant, 2)
        s[:32], s[32:64] = a, b
        res = g.stcon_batch(s, t)
        say(f"rmat24 batched: {n} queries in {res.stats['time_ms']:.1f} ms = "
            f"{n / (res.stats['time_ms'] / 1e3):.0f} q/s; unconnected {(res.dist < 0).sum()}")
        assert ((res.dist < 0) == far)[64:].mean() > 0.99  # giant-giant pairs connect, the rest do not
        idx = np.concatenate([np.arange(64), rng.choice(np.arange(64, n), 960, replace=False)])
        plain = g.stcon_many(s[idx], t[idx])
        assert np.array_equal(plain.dist, res.dist[idx])
        for src, sl in ((a, slice(0, 32)), (b, slice(32, 64))):
            d, _, _ = orc.bfs(hg.offsets, hg.columns, int(src))
            assert np.array_equal(res.dist[sl], d[t[sl]])
