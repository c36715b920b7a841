import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "fast-con_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def fc():
    import fastcon

    return fastcon


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle

    return oracle


def small_graphs(fc):
    """A family of small graphs with known structure, as (name, HostGraph)."""
    out = []
    # path 0-1-...-9
    e = np.array([[i, i + 1] for i in range(9)])
    out.append(("path10", fc.csr_from_edges(e, 10)))
    # star centred at 0 with 20 leaves, plus an isolated vertex 21
    e = np.array([[0, i] for i in range(1, 21)])
    out.append(("star20+iso", fc.csr_from_edges(e, 22)))
    # cycle of 13
    e = np.array([[i, (i + 1) % 13] for i in range(13)])
    out.append(("cycle13", fc.csr_from_edges(e, 13)))
    # complete graph K7
    e = np.array([[i, j] for i in range(7) for j in range(i + 1, 7)])
    out.append(("K7", fc.csr_from_edges(e, 7)))
    # two components
    e = np.array([[0, 1], [1, 2], [3, 4], [4, 5], [5, 3]])
    out.append(("two-comp", fc.csr_from_edges(e, 8)))
    out.append(("grid9x7", fc.gen_grid(9, 7, 0.0, 1)))
    out.append(("grid40x40-p10", fc.gen_grid(40, 40, 0.10, 7)))
    out.append(("er-500-d3", fc.gen_er(500, 3.0, 11)))
    out.append(("er-3000-d1.2", fc.gen_er(3000, 1.2, 12)))
    out.append(("rmat10", fc.gen_rmat(10, 16, seed=5)))
    out.append(("rmat12-skew", fc.gen_rmat(12, 8, a=0.65, b=0.15, c=0.15, seed=6)))
    return out


def oracle_pair_dists(orc, hg, s, t, workers=16, chunk=8):
    """Host-BFS s–t distances (BFS from s[i], level of t[i]) for many pairs,
    in a thread pool: the oracle's C calls release the GIL."""
    from concurrent.futures import ThreadPoolExecutor

    s = np.ascontiguousarray(s, np.int32)
    t = np.ascontiguousarray(t, np.int32)
    out = np.empty(len(s), np.int32)
    workers = max(1, min(workers, os.cpu_count() or 1))

    def run(k):
        _, d = orc.time_stcon(hg.offsets, hg.columns, s[k:k + chunk], t[k:k + chunk])
        out[k:k + chunk] = d

    with ThreadPoolExecutor(workers) as ex:
        list(ex.map(run, range(0, len(s), chunk)))
    return out


def oracle_bfs_many(orc, hg, roots, fn, workers=16):
    """fn(root, dist, info) for the host BFS of every root, in a thread pool;
    returns the list of fn's results in root order."""
    from concurrent.futures import ThreadPoolExecutor

    workers = max(1, min(workers, os.cpu_count() or 1))

    def run(r):
        d, _, info = orc.bfs(hg.offsets, hg.columns, int(r))
        return fn(int(r), d, info)

    with ThreadPoolExecutor(workers) as ex:
        return list(ex.map(run, roots))
