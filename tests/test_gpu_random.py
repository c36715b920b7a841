"""Seeded sweep over small graphs of assorted structure: random trees,
cliques joined by long paths, hub-and-spoke graphs, unions of components,
sparse ER graphs and small skewed R-MATs. On each, BFS levels and parents,
the plain s–t path, the batched engine under the plain, heuristic and
learned schedules (the learned one with random network weights and no
margin, so that it takes arbitrary actions, bottom-up and wide steps
included) and the helper-seed verdict are compared with the host BFS. Half
of the graphs are built with a low hub threshold so that the hub queues and
hub lists carry ordinary vertices too; every third graph is uploaded with
64-bit offsets and every fourth keeps the caller's vertex numbering. The
plain visit's virtual-warp width is fixed to 4, 8, 16, 32 or 64 on every
other graph, and the grid varies between 1 and 4 blocks per CU. One more
batched run per graph draws its width, hub chunk, lane count, launch mode
and blocks per CU at random."""
import os
import sys

import numpy as np
import pytest

from oracle import mlp

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _tree(rng, n):
    par = np.array([rng.integers(0, i) for i in range(1, n)])
    return np.stack([np.arange(1, n), par], 1), n


def _cliques_on_a_path(rng, k, size, gap):
    edges, base = [], 0
    prev_tail = None
    for _ in range(k):
        members = np.arange(base, base + size)
        edges += [[a, b] for i, a in enumerate(members) for b in members[i + 1:]]
        base += size
        if prev_tail is not None:  # a path of `gap` new vertices from the previous clique
            chain = [prev_tail] + list(range(base, base + gap)) + [members[0]]
            edges += [[chain[i], chain[i + 1]] for i in range(len(chain) - 1)]
            base += gap
        prev_tail = members[-1]
    return np.array(edges), base


def _hubs(rng, n, nhub, p_hub):
    edges = [[h, v] for h in range(nhub) for v in range(nhub, n) if rng.random() < p_hub]
    edges += [[h, g] for h in range(nhub) for g in range(h + 1, nhub)]
    edges += [[v, v + 1] for v in range(nhub, n - 1, 7)]
    return np.array(edges), n


def _components(rng, sizes):
    edges, base = [], 0
    for sz in sizes:
        m = max(1, int(sz * 1.5))
        a = rng.integers(0, sz, m) + base
        b = rng.integers(0, sz, m) + base
        edges += np.stack([a, b], 1).tolist()
        base += sz
    return np.array(edges), base + 3  # three isolated vertices at the end


def _family(fc, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(4):
        e, n = _tree(rng, int(rng.integers(50, 3000)))
        out.append((f"tree{i}", fc.csr_from_edges(e, n)))
    for i in range(3):
        e, n = _cliques_on_a_path(rng, int(rng.integers(2, 6)), int(rng.integers(5, 40)), int(rng.integers(1, 60)))
        out.append((f"cliques{i}", fc.csr_from_edges(e, n)))
    for i in range(3):
        e, n = _hubs(rng, int(rng.integers(300, 3000)), int(rng.integers(2, 9)), float(rng.uniform(0.05, 0.6)))
        out.append((f"hubs{i}", fc.csr_from_edges(e, n)))
    for i in range(3):
        e, n = _components(rng, [int(x) for x in rng.integers(2, 800, int(rng.integers(2, 9)))])
        out.append((f"components{i}", fc.csr_from_edges(e, n)))
    for i in range(3):
        out.append((f"er{i}", fc.gen_er(int(rng.integers(200, 6000)), float(rng.uniform(0.8, 6.0)),
                                        int(rng.integers(1, 1 << 30)))))
    for i in range(3):
        a = float(rng.uniform(0.45, 0.7))
        bc = (1.0 - a - 0.05) / 2
        out.append((f"rmat{i}", fc.gen_rmat(int(rng.integers(8, 13)), float(rng.choice([4.0, 8.0, 16.0])), a=a,
                                            b=bc, c=bc, seed=int(rng.integers(1, 1 << 30)))))
    return out


def _pairs(rng, hg, n):
    s = rng.integers(0, hg.V, n).astype(np.int32)
    t = rng.integers(0, hg.V, n).astype(np.int32)
    s[:2] = t[:2]
    return s, t


@pytest.mark.parametrize("seed", [1, 2, 3, 4, 5, 6])
def test_random_family_matches_host_bfs(fc, orc, monkeypatch, seed):
    rng = np.random.default_rng(100 + seed)
    for k, (name, hg) in enumerate(_family(fc, seed)):
        low_hub = k % 2 == 1
        if low_hub:
            monkeypatch.setenv("FASTCON_HUB_THRESHOLD", "6")
        else:
            monkeypatch.delenv("FASTCON_HUB_THRESHOLD", raising=False)
        vw = (4, 8, 16, 32, 64)[(k // 2) % 5] if k % 2 == 0 else 0
        if vw:
            monkeypatch.setenv("FASTCON_VW", str(vw))
        else:
            monkeypatch.delenv("FASTCON_VW", raising=False)
        monkeypatch.setenv("FASTCON_BLOCKS_PER_CU", str(1 + k % 4))
        monkeypatch.setenv("FASTCON_BATCH_BLOCKS_PER_CU", str(1 + k % 3))
        s, t = _pairs(rng, hg, 192)
        exp = np.array([orc.stcon(hg.offsets, hg.columns, a, b)[1] for a, b in zip(s, t)], np.int32)
        roots = rng.integers(0, hg.V, 3)
        off64, keep = k % 3 == 2, k % 4 == 3
        with fc.Graph(hg, force64=off64, keep_ids=keep) as g:
            for r in roots:
                say(f"random seed={seed} {name} V={hg.V} E={hg.E} low_hub={low_hub} off64={off64} keep_ids={keep} "
                    f"vw={vw} bpc={1 + k % 4} bfs root={r}")
                lv, par, _ = g.bfs(int(r), levels=True, parents=True)
                assert np.array_equal(lv, orc.bfs(hg.offsets, hg.columns, int(r))[0]), name
                assert orc.validate_bfs_tree(hg.offsets, hg.columns, int(r), lv, par) == 0, name
            say(f"random seed={seed} {name} plain s-t")
            assert np.array_equal(g.stcon_many(s, t).dist, exp), name
            for policy in (0, 1):
                say(f"random seed={seed} {name} batch policy={policy}")
                res = g.stcon_batch(s, t, policy=policy)
                bad = np.flatnonzero(res.dist != exp)
                assert bad.size == 0, (name, policy, bad[:5], res.dist[bad[:5]], exp[bad[:5]])
            bvw = int(rng.choice([0, 4, 8, 16, 32, 64]))
            btile = int(rng.choice([64, 128, 256, 512, 1024, 4096]))
            blanes = int(rng.integers(1, 5))
            bgraph = int(rng.integers(0, 2))
            # bridge steps before every eligible step, before large ones only,
            # or (the default threshold) hardly ever on graphs this small
            bmin = int(rng.choice([1, 64, 1 << 20]))
            old_bmin = os.environ.get("FASTCON_BRIDGE_MIN_EDGES")
            os.environ["FASTCON_BRIDGE_MIN_EDGES"] = str(bmin)
            try:
                say(f"random seed={seed} {name} batch vw={bvw} tile={btile} lanes={blanes} graph={bgraph} "
                    f"bpc={1 + k % 3} bridge_min={bmin}")
                res = g.stcon_batch(s, t, policy=fc.POLICY_HEURISTIC, vw=bvw, tile=btile, use_graph=bgraph,
                                    lanes=blanes)
                bad = np.flatnonzero(res.dist != exp)
                assert bad.size == 0, (name, bvw, btile, blanes, bgraph, bmin, bad[:5], res.dist[bad[:5]],
                                       exp[bad[:5]])
                w = rng.normal(0, 0.5, mlp.P).astype(np.float32)
                g.sched_set_params(w)
                for lanes in (1, 3):
                    say(f"random seed={seed} {name} batch learned (random weights) lanes={lanes} bridge_min={bmin}")
                    res = g.stcon_batch(s, t, policy=fc.POLICY_LEARNED, learn_margin=0.0, lanes=lanes)
                    bad = np.flatnonzero(res.dist != exp)
                    assert bad.size == 0, (name, "learned", lanes, bmin, bad[:5], res.dist[bad[:5]], exp[bad[:5]])
            finally:
                if old_bmin is None:
                    os.environ.pop("FASTCON_BRIDGE_MIN_EDGES", None)
                else:
                    os.environ["FASTCON_BRIDGE_MIN_EDGES"] = old_bmin
            for a, b, d in list(zip(s, t, exp))[2:6]:
                say(f"random seed={seed} {name} seeds s={a} t={b}")
                assert g.stcon_seeds(int(a), int(b), helpers=int(rng.integers(1, 64)),
                                     seed=int(rng.integers(1, 1000)))[0] == int(d >= 0), name
