"""GPU parity on the benchmark workload itself: undirected R-MAT scale 22
(edgefactor 16, a=0.57 b=0.19 c=0.19, seed 22), with the benchmark's 64 BFS
roots and 1,024 (s, t) pairs (fc_sample_roots, seed 23).

* BFS from each of the 64 roots: levels equal the host BFS's
  (BFS-4K/src/host/graph.cpp:83-97) bit for bit, the parent array is a
  valid BFS tree (the Graph500 rules of
  st-conn/reference_functions.c:99-434, checked by the oracle), and the
  visited-edge count equals Graph::visitedEdges
  (BFS-4K/src/host/graph.cpp:103-110), as cudaGraph's Compare does for the
  levels (BFS-4K/src/device/BFS_WorkEfficient.cu:47-51).
* The 1,024 pairs: verdicts and distances of the plain one-query-per-launch
  visit (fc_stcon_many) and of the batched engine (fc_stcon_batch, under
  the plain, heuristic and shipped learned schedules) equal the host BFS's
  level of t from s, for every pair.

The host BFSs run in a thread pool (about 50 s of CPU work on 16 cores)."""
import sys
import time

import numpy as np
import pytest

from conftest import oracle_bfs_many, oracle_pair_dists

pytestmark = pytest.mark.gpu

SCALE, EF, SEED, ROOTS, PAIRS = 22, 16, 22, 64, 1024


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def s22(fc):
    t0 = time.time()
    hg = fc.gen_rmat(SCALE, EF, 0.57, 0.19, 0.19, seed=SEED)
    roots = hg.sample_roots(ROOTS + 2 * PAIRS, SEED + 1)
    say(f"rmat22: V={hg.V} E={hg.E} built in {time.time() - t0:.1f}s")
    return hg, roots[:ROOTS], roots[ROOTS:ROOTS + PAIRS], roots[ROOTS + PAIRS:]


@pytest.mark.timeout(600)
def test_s22_bfs_levels_and_trees(fc, orc, s22):
    hg, roots, _, _ = s22
    t0 = time.time()
    with fc.Graph(hg) as g:
        for k in range(0, ROOTS, 16):
            chunk = roots[k:k + 16]
            got = []
            for r in chunk:
                lv, pa, st = g.bfs(int(r), parents=True)
                got.append((lv, pa, st))
            by_root = {int(r): x for r, x in zip(chunk, got)}

            def check(r, d, info):
                lv, pa, st = by_root[r]
                ok_lv = np.array_equal(lv, d)
                ok_tree = orc.validate_bfs_tree(hg.offsets, hg.columns, r, lv, pa) == 0
                ok_cnt = (st["visited_vertices"], st["visited_edges"], st["levels"] - 1) == info
                return ok_lv, ok_tree, ok_cnt

            res = oracle_bfs_many(orc, hg, chunk, check)
            for r, (ok_lv, ok_tree, ok_cnt) in zip(chunk, res):
                assert ok_lv, f"BFS levels from root {r} differ from the host BFS"
                assert ok_tree, f"BFS parents from root {r} are not a valid BFS tree"
                assert ok_cnt, f"BFS counts from root {r} differ from visitedNodes/visitedEdges/getMaxDistance"
            say(f"rmat22 BFS roots {k}..{k + len(chunk) - 1} checked ({time.time() - t0:.1f}s)")


@pytest.mark.timeout(900)
def test_s22_pairs_plain_and_batched(fc, orc, s22):
    hg, _, s, t = s22
    t0 = time.time()
    exp = oracle_pair_dists(orc, hg, s, t)
    say(f"rmat22: host BFS for {PAIRS} pairs in {time.time() - t0:.1f}s; distances {exp.min()}..{exp.max()}, "
        f"unconnected {(exp < 0).sum()}")
    with fc.Graph(hg) as g:
        plain = g.stcon_many(s, t)
        assert np.array_equal(plain.dist, exp), "plain s–t distances differ from the host BFS"
        assert np.array_equal(plain.connected, (exp >= 0).astype(np.int32))
        say(f"rmat22 plain: {PAIRS / (plain.stats['time_ms'] / 1e3):.0f} q/s, all {PAIRS} pairs match")
        pols = [("plain", fc.POLICY_PLAIN), ("heuristic", fc.POLICY_HEURISTIC)]
        if fc.load_default_sched_params(g):
            pols.append(("learned", fc.POLICY_LEARNED))
        for name, pol in pols:
            res = g.stcon_batch(s, t, policy=pol)
            assert np.array_equal(res.dist, exp), f"batched ({name}) distances differ from the host BFS"
            assert np.array_equal(res.connected, (exp >= 0).astype(np.int32))
            say(f"rmat22 batched {name}: {PAIRS / (res.stats['time_ms'] / 1e3):.0f} q/s, all {PAIRS} pairs match")
