"""The bfs4K and st-con command lines: argument handling as the reference's,
their printed numbers and the st_stats file checked against the host BFS."""
import glob
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "fast-con_amd", "bin")


def _undirected_pairs(hg):
    src = np.repeat(np.arange(hg.V), np.diff(hg.offsets))
    keep = src < hg.columns
    return src[keep], hg.columns[keep]


def _write_mm(path, hg):
    u, v = _undirected_pairs(hg)
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate pattern symmetric\n")
        f.write(f"{hg.V} {hg.V} {len(u)}\n")
        for a, b in zip(u, v):
            f.write(f"{a + 1} {b + 1}\n")


def _write_edge_list(path, hg):
    u, v = _undirected_pairs(hg)
    with open(path, "w") as f:
        f.write(f"# Nodes: {hg.V} Edges: {len(u)}\n# FromNodeId\tToNodeId\n")
        for a, b in zip(u, v):
            f.write(f"{a}\t{b}\n")


def _run(args, env=None):
    e = dict(os.environ)
    if env:
        e.update(env)
    return subprocess.run(args, capture_output=True, text=True, timeout=300, env=e)


def test_bfs4k_matches_host_bfs(fc, orc, tmp_path):
    hg = fc.gen_rmat(11, 8, seed=3)
    path = str(tmp_path / "g.mtx")
    _write_mm(path, hg)
    r = _run([os.path.join(BIN, "bfs4K"), path, "-n", "1"])
    assert r.returncode == 0, r.stderr
    d, _, info = orc.bfs(hg.offsets, hg.columns, 0)
    assert re.search(r"Number of TESTS:  1\n", r.stdout)
    assert int(re.search(r"Avg. Diameter:  (\d+)", r.stdout).group(1)) == d.max()
    assert float(re.search(r"Avg. MTEPS:  ([0-9.]+)", r.stdout).group(1)) > 0
    assert re.search(r"Nodes: 2,048\tEdges: [0-9,]+\tGraphType: Undirected \(File Def\)", r.stdout)

    r = _run([os.path.join(BIN, "bfs4K"), path, "-U", "-n", "5"], env={"FASTCON_SEED": "7"})
    assert r.returncode == 0, r.stderr
    iters = re.findall(r"iter: (\d+)\s+time: \S+\s+Edges: (\d+)\s+source: (\d+)", r.stdout)
    assert len(iters) == 5
    ecc = []
    for _, edges, src in iters:
        d, _, info = orc.bfs(hg.offsets, hg.columns, int(src))
        assert int(edges) == info[1]
        ecc.append(int(d.max()))
    assert int(re.search(r"Avg. Diameter:  (\d+)", r.stdout).group(1)) == sum(ecc) // 5


def test_bfs4k_rejects_bad_arguments(tmp_path):
    r = _run([os.path.join(BIN, "bfs4K")])
    assert r.returncode != 0 and "Syntax Error" in r.stderr
    p = tmp_path / "x.mtx"
    p.write_text("%%MatrixMarket matrix coordinate pattern symmetric\n2 2 1\n1 2\n")
    for bad in (["-x"], ["-n"], ["-n", "3a"]):
        r = _run([os.path.join(BIN, "bfs4K"), str(p)] + bad)
        assert r.returncode != 0 and "Syntax Error" in r.stderr, bad


def test_stcon_cli_stats_file(fc, orc, tmp_path):
    hg = fc.gen_rmat(11, 8, seed=5)
    path = str(tmp_path / "g.txt")
    _write_edge_list(path, hg)
    r = _run([os.path.join(BIN, "st-con"), "-f", path, "-N", "12", "-2", "9", "-d", str(tmp_path)])
    assert r.returncode == 0, r.stderr
    files = glob.glob(str(tmp_path / "st_stats_*"))
    assert len(files) == 1
    text = open(files[0]).read().splitlines()
    assert text[0].startswith("SCALE ") and text[3] == "Num-ST-CON=12" and text[4] == f"Filename {path}"
    head = text.index("ST-CON-RUN RED-V BLUE-V DEPTH MN-FOUND TIME(Sec) TEPS")
    rows = [line.split() for line in text[head + 1:head + 13]]
    for k, row in enumerate(rows):
        i, s, t, depth, found = map(int, row[:5])
        assert i == k
        conn, d, dp = orc.stcon(hg.offsets, hg.columns, s, t)
        assert found == int(conn)
        if conn:
            assert depth == dp == (d + 1) // 2
    assert text[head + 13].startswith("max-ST-CON-level ")
    assert r.stdout.count("END STCON [") == 12
    # validation is on by default, as the reference's -V 1: one OK per
    # connected pair with s != t
    nconn = sum(1 for row in rows if int(row[4]) and row[1] != row[2])
    assert r.stdout.count("ST-CON Validation 0 OK with Matching node") == nconn and "ERROR" not in r.stdout
    r = _run([os.path.join(BIN, "st-con"), "-f", path, "-N", "4", "-2", "9", "-V", "0", "-d", str(tmp_path)])
    assert r.returncode == 0 and "ST-CON Validation" not in r.stdout
    r = _run([os.path.join(BIN, "st-con"), "-v"])
    assert r.returncode == 0 and r.stdout.startswith("Version ")
    r = _run([os.path.join(BIN, "st-con")])
    assert r.returncode != 0 and "Usage" in r.stderr
