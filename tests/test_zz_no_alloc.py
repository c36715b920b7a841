"""fc_stcon_many_static, the entry point that allocates no device memory,
against fc_stcon_many on the same graphs and pairs, through the C ABI.

The CSR graphs are put in device memory by the test itself (hipMalloc and
hipMemcpy of the HIP runtime, called through ctypes), as a caller that owns
its device memory would. Two calls in a row: first R-MAT s16 (hub vertices,
1,500 queries, so two passes of the fixed query list), then a smaller
sparse Erdős–Rényi graph with many unconnected pairs and fewer queries, so
the second call runs over a workspace the first left behind. Distances and
verdicts must equal fc_stcon_many's bit for bit, and the host BFS's on a
sample."""
import ctypes as C
import ctypes.util

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Hip:
    def __init__(self):
        name = ctypes.util.find_library("amdhip64") or "libamdhip64.so"
        self.lib = C.CDLL(name)
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), max(a.nbytes, 4)) == 0
        self.bufs.append(p)
        if a.nbytes:
            assert self.lib.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0  # host to device
        return p.value

    def free(self):
        for p in self.bufs:
            self.lib.hipFree(p)
        self.bufs = []


def _check_static(fc, orc, hip, name, hg, s, t):
    off = hg.offsets.astype(np.int32)
    d_off, d_col = hip.put(off), hip.put(hg.columns)
    got = fc.stcon_many_static(d_off, d_col, hg.V, hg.E, s, t)
    with fc.Graph(hg) as g:
        ref = g.stcon_many(s, t)
    assert np.array_equal(got.dist, ref.dist), f"{name}: static and fc_stcon_many distances differ"
    assert np.array_equal(got.connected, ref.connected), f"{name}: verdicts differ"
    for i in range(0, len(s), max(1, len(s) // 12)):
        c, d, _ = orc.stcon(hg.offsets, hg.columns, int(s[i]), int(t[i]))
        assert got.dist[i] == d and got.connected[i] == int(c), f"{name}: pair {i} differs from the host BFS"
    print(f"{name}: {len(s)} queries, {int((got.dist < 0).sum())} unconnected, "
          f"{got.stats['launches']} launches, match", flush=True)
    return got


def test_static_entry_point_matches_fc_stcon_many(fc, orc):
    hip = _Hip()
    try:
        big = fc.gen_rmat(16, 16, seed=61)
        assert (big.degree() >= 512).any()  # hub-queue path
        r = big.sample_roots(3000, 62)
        s, t = r[:1500].copy(), r[1500:].copy()
        s[:7] = t[:7]  # s == t queries
        _check_static(fc, orc, hip, "rmat16", big, s, t)
        small = fc.gen_er(3000, 1.2, 12)
        rng = np.random.default_rng(63)
        s2 = rng.integers(0, small.V, 200).astype(np.int32)
        t2 = rng.integers(0, small.V, 200).astype(np.int32)
        res = _check_static(fc, orc, hip, "er3000", small, s2, t2)
        assert (res.dist < 0).any() and (res.dist > 0).any()
    finally:
        hip.free()


def test_static_entry_point_limits(fc):
    with pytest.raises(fc.FastconError) as ei:
        fc.stcon_many_static(1, 1, (1 << 20) + 1, 10, [0], [1])
    assert ei.value.status == fc.ERR_INVALID
    with pytest.raises(fc.FastconError) as ei:
        fc.stcon_many_static(1, 1, 100, (1 << 25) + 1, [0], [1])
    assert ei.value.status == fc.ERR_INVALID
    with pytest.raises(fc.FastconError) as ei:
        fc.stcon_many_static(1, 1, 100, 10, [0], [100])
    assert ei.value.status == fc.ERR_INVALID
