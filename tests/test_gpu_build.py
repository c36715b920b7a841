"""GPU graph construction: fc_csr_from_edges_gpu and fc_gen_rmat_gpu give the
same CSR arrays, bit for bit, as the host canonicaliser and generator (which
the CPU tests pin against the oracle), on edge lists with self-loops,
dropped (-1) pairs, duplicates and isolated vertices."""
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _same(a, b):
    return a.V == b.V and a.E == b.E and np.array_equal(a.offsets, b.offsets) and np.array_equal(a.columns, b.columns)


def test_csr_from_edges_gpu_equals_host(fc, orc):
    rng = np.random.default_rng(5)
    cases = [(1, 0), (5, 3), (7, 50), (100, 1000), (2000, 30000), (70000, 300000), (1 << 18, 1 << 21)]
    for V, m in cases:
        e = rng.integers(-1, V, size=(m, 2)) if m else np.zeros((0, 2), np.int64)
        if m > 10:
            e[: m // 10, 1] = e[: m // 10, 0]          # self-loops
            e[m // 10: m // 5] = e[m // 5: m // 5 + (m // 5 - m // 10)]  # duplicates
        say(f"edges V={V} m={m}")
        host = fc.csr_from_edges(e, V)
        gpu = fc.csr_from_edges_gpu(e, V)
        assert _same(host, gpu), (V, m)
        if m:
            assert _same(fc.csr_from_edges(e, 0), fc.csr_from_edges_gpu(e, 0))
    off, col = orc.canonical(e, V)
    assert np.array_equal(gpu.offsets, off) and np.array_equal(gpu.columns, col)
    with pytest.raises(fc.FastconError) as ei:
        fc.csr_from_edges_gpu(np.array([[0, 5]]), 5)
    assert ei.value.status == fc.ERR_INVALID


@pytest.mark.parametrize("scale,ef,a,seed", [(8, 16, 0.57, 1), (14, 8, 0.45, 2), (18, 16, 0.57, 3), (20, 16, 0.65, 4)])
def test_gen_rmat_gpu_equals_host(fc, scale, ef, a, seed):
    b = (1.0 - a - 0.05) / 2
    t0 = time.time()
    host = fc.gen_rmat(scale, ef, a, b, b, seed=seed)
    t1 = time.time()
    gpu = fc.gen_rmat_gpu(scale, ef, a, b, b, seed=seed)
    t2 = time.time()
    say(f"rmat s{scale} ef{ef}: V={host.V} E={host.E}; host {t1 - t0:.3f}s, gpu {t2 - t1:.3f}s")
    assert _same(host, gpu)


def test_csr_from_edges_gpu_edge_shapes(fc, orc):
    """Shapes the radix build must get right: nothing but self-loops and
    dropped pairs, one edge repeated many times, a star whose hub has 200 K
    neighbours, ids at the top of the range, and vertex counts on either side
    of a power of two (the key width changes there)."""
    rng = np.random.default_rng(9)
    cases = {
        "self-loops and dropped pairs only": (50, np.array([[i, i] for i in range(50)] + [[-1, -1]] * 10)),
        "one edge repeated": (10, np.array([[3, 7]] * 5000)),
        "star of 200K leaves": (200001, np.stack([np.zeros(200000, np.int64), np.arange(1, 200001)], 1)),
        "top ids": (1000, np.array([[999, 998], [998, 999], [999, 0], [0, 999], [997, 999]])),
    }
    for V in ((1 << 16), (1 << 16) + 1, (1 << 20) - 1, 1 << 20):
        e = rng.integers(0, V, size=(4096, 2))
        e[:16, 0] = V - 1
        cases[f"V={V}"] = (V, e)
    for name, (V, e) in cases.items():
        say(f"edge shape: {name}")
        host = fc.csr_from_edges(e, V)
        gpu = fc.csr_from_edges_gpu(e, V)
        assert _same(host, gpu), name
        off, col = orc.canonical(e, V)
        assert np.array_equal(gpu.offsets, off) and np.array_equal(gpu.columns, col), name
