"""GPU parity on the learned-scheduler graph family: one graph of each kind,
4,096 seeded s–t queries per graph.

* R-MAT scale 20 with a = 0.65 and scale 24 with a = 0.45 (b = c =
  (1 - a - 0.05) / 2, edgefactor 16; the scale-24 graph is built on the GPU,
  bit for bit the host builder's graph);
* Erdős–Rényi G(n, M) with n = 8,388,608 and mean degree 8;
* the 2048 x 2048 grid (no deletions, so every distance is the Manhattan
  distance: a closed form for all 4,096 queries).

For each graph the batched engine under the shipped learned schedule (the
heuristic when none ships) must give the plain visit's verdict and distance
for every query, and the host BFS's for the first 256 queries, which share
eight sources so that eight host BFSs check them. The batch times of the
plain, heuristic and learned schedules are printed with the learned
schedule's speed-up over the other two (the scheduler reorders work only, so
its speed-up is the whole of its effect)."""
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 4096


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _queries(hg, seed):
    r = hg.sample_roots(2 * N, seed)
    s, t = r[:N].copy(), r[N:].copy()
    src = r[:8]
    s[:256] = np.repeat(src, 32)
    return s, t, src


def _family_check(fc, orc, name, hg, seed, closed_form=None, plain_subset=None):
    s, t, src = _queries(hg, seed)
    exp = np.empty(256, np.int32)
    for k, r in enumerate(src):
        d, _, _ = orc.bfs(hg.offsets, hg.columns, int(r))
        exp[32 * k:32 * (k + 1)] = d[t[32 * k:32 * (k + 1)]]
    with fc.Graph(hg) as g:
        learned = fc.load_default_sched_params(g)
        pol = fc.POLICY_LEARNED if learned else fc.POLICY_HEURISTIC
        res = g.stcon_batch(s, t, policy=pol)
        assert np.array_equal(res.dist[:256], exp), f"{name}: batched distances differ from the host BFS"
        if closed_form is not None:
            assert np.array_equal(res.dist, closed_form(s, t)), f"{name}: batched distances differ from the closed form"
        idx = np.arange(N) if plain_subset is None else np.arange(plain_subset)
        plain = g.stcon_many(s[idx], t[idx])
        assert np.array_equal(plain.dist, res.dist[idx]), f"{name}: batched and plain distances differ"
        assert np.array_equal(plain.dist[:256], exp), f"{name}: plain distances differ from the host BFS"
        ms = {}
        for pname, p in (("plain", fc.POLICY_PLAIN), ("heuristic", fc.POLICY_HEURISTIC),
                         ("learned", fc.POLICY_LEARNED)):
            if p == fc.POLICY_LEARNED and not learned:
                continue
            times = []
            for _ in range(3):
                r2 = g.stcon_batch(s, t, policy=p)
                assert np.array_equal(r2.dist, res.dist), f"{name}: schedule {pname} changed an answer"
                times.append(r2.stats["time_ms"])
            ms[pname] = float(np.median(times))
        pq = len(idx) / (plain.stats["time_ms"] / 1e3)
        line = (f"{name}: V={hg.V} E={hg.E}; plain visit {pq:.0f} q/s; batch ms " +
                ", ".join(f"{k} {v:.2f}" for k, v in ms.items()))
        if "learned" in ms:
            line += (f"; learned/heuristic speed-up {ms['heuristic'] / ms['learned']:.3f}x, "
                     f"learned/plain {ms['plain'] / ms['learned']:.2f}x")
        say(line)
        return ms


@pytest.mark.timeout(300)
def test_family_rmat20_a065(fc, orc):
    hg = fc.gen_rmat(20, 16, 0.65, 0.15, 0.15, seed=2065)
    _family_check(fc, orc, "rmat20 a=0.65", hg, 2066)


@pytest.mark.timeout(600)
def test_family_rmat24_a045(fc, orc):
    t0 = time.time()
    hg = fc.gen_rmat_gpu(24, 16, 0.45, 0.25, 0.25, seed=2445)
    say(f"rmat24 a=0.45 built on the GPU in {time.time() - t0:.1f}s")
    _family_check(fc, orc, "rmat24 a=0.45", hg, 2446)


@pytest.mark.timeout(300)
def test_family_er8m_d8(fc, orc):
    hg = fc.gen_er(8 * 2**20, 8.0, seed=808)
    _family_check(fc, orc, "ER 8M d=8", hg, 809)


@pytest.mark.timeout(600)
def test_family_grid2048(fc, orc):
    n = 2048
    hg = fc.gen_grid(n, n, 0.0, 2048)

    def manhattan(s, t):
        return (np.abs(s // n - t // n) + np.abs(s % n - t % n)).astype(np.int32)

    _family_check(fc, orc, "grid 2048^2", hg, 2049, closed_form=manhattan, plain_subset=1024)
