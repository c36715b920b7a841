"""Writes tests/golden/generators.json: for each seeded graph the tests and
the benchmark use, its vertex and directed-edge counts, isolated vertices,
largest degree, the histogram of floor(log2(degree)) over vertices of
degree > 0, and SHA-256 digests of the CSR arrays (offsets as little-endian
int64, columns as little-endian int32). The generators are this engine's own
(fc_gen_rmat, fc_gen_grid, fc_gen_er); the fixture pins their output so that
a change to any of them (or to the canonicalisation) shows up as a test
failure rather than as a silently different benchmark graph.

    python tests/golden/make_generator_fixtures.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "fast-con_amd")]

CONFIGS = [
    {"name": "rmat16", "kind": "rmat", "scale": 16, "edgefactor": 16, "abc": [0.57, 0.19, 0.19], "seed": 16},
    {"name": "rmat22", "kind": "rmat", "scale": 22, "edgefactor": 16, "abc": [0.57, 0.19, 0.19], "seed": 22},
    {"name": "grid4096_p002", "kind": "grid", "rows": 4096, "cols": 4096, "p_delete": 0.02, "seed": 7},
    {"name": "er1m_d8", "kind": "er", "n": 1 << 20, "mean_degree": 8.0, "seed": 808},
]


def build(fc, c):
    if c["kind"] == "rmat":
        a, b, cc = c["abc"]
        return fc.gen_rmat(c["scale"], c["edgefactor"], a, b, cc, seed=c["seed"])
    if c["kind"] == "grid":
        return fc.gen_grid(c["rows"], c["cols"], c["p_delete"], c["seed"])
    return fc.gen_er(c["n"], c["mean_degree"], seed=c["seed"])


def describe(hg):
    deg = hg.degree()
    pos = deg[deg > 0]
    hist = np.bincount(np.floor(np.log2(pos)).astype(np.int64)) if len(pos) else np.zeros(0, np.int64)
    return {
        "V": int(hg.V),
        "E": int(hg.E),
        "isolated": int((deg == 0).sum()),
        "max_degree": int(deg.max()) if len(deg) else 0,
        "log2_degree_hist": [int(x) for x in hist],
        "sha256_offsets": hashlib.sha256(hg.offsets.astype("<i8").tobytes()).hexdigest(),
        "sha256_columns": hashlib.sha256(hg.columns.astype("<i4").tobytes()).hexdigest(),
    }


def main():
    import fastcon as fc

    out = []
    for c in CONFIGS:
        d = dict(c)
        d.update(describe(build(fc, c)))
        out.append(d)
        print(c["name"], d["V"], d["E"], flush=True)
    with open(os.path.join(HERE, "generators.json"), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
