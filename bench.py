"""Benchmark of the BFS / s–t connectivity hot path on one MI355X.

Workload (one "step"): 1,024 seeded s–t queries on an undirected R-MAT
scale-22 graph (edgefactor 16, a=0.57 b=0.19 c=0.19 d=0.05, int32 CSR),
answered by the engine's fastest exact path. Inputs are resident in HBM
before the timed region. Also measured in the same run: the plain
one-query-per-launch path on the same pairs, and BFS from 64 seeded roots
(GTEPS). The CPU baseline is the oracle's single-thread host BFS s–t
(BFS from s, read the level of t) on a bounded sample of the same pairs.

Prints one JSON line.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "fast-con_amd"))

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edgefactor", type=float, default=16.0)
    ap.add_argument("--pairs", type=int, default=1024)
    ap.add_argument("--roots", type=int, default=64)
    ap.add_argument("--seed", type=int, default=22)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="bound on the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["all", "plain", "bfs"], default="all")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import fastcon as fc

    t0 = time.time()
    hg = fc.gen_rmat(args.scale, args.edgefactor, 0.57, 0.19, 0.19, seed=args.seed)
    V, E = hg.V, hg.E
    roots = hg.sample_roots(args.roots + 2 * args.pairs, args.seed + 1)
    bfs_roots = roots[: args.roots]
    s = roots[args.roots: args.roots + args.pairs].copy()
    t = roots[args.roots + args.pairs:].copy()
    log(f"graph: R-MAT s{args.scale} V={V} E={E} built in {time.time() - t0:.1f}s")
    g = fc.Graph(hg)
    log(f"uploaded: offsets {'int64' if g.offsets64 else 'int32'}")

    result = {}
    # ---- headline: s–t queries/s over the pairs, K timed steps ----
    def plain_step():
        return g.stcon_many(s, t, policy=fc.POLICY_PLAIN)

    step_fn = plain_step
    path = "plain"
    ref = step_fn()  # also the first warm-up step
    for _ in range(args.warmup - 1):
        step_fn()
    ms = []
    dev_ms = []
    stats = None
    for _ in range(args.steps):
        a = time.perf_counter()
        r = step_fn()
        ms.append((time.perf_counter() - a) * 1e3)
        dev_ms.append(r.stats["time_ms"])
        stats = r.stats
        assert np.array_equal(r.dist, ref.dist)
    ms_step = float(np.mean(ms))
    qps = args.pairs / (ms_step / 1e3)
    plain_qps = qps
    log(f"{path}: {ms_step:.2f} ms/step wall, {np.mean(dev_ms):.2f} ms device, {qps:.0f} q/s")

    # Roofline of the dominant kernel (the per-query visit launch):
    # B_q = 8*E_x + 20*V_x + 4*V algorithmic bytes per launch, E_x and V_x the
    # expanded edges and vertices reported by the kernel.
    launches = int(np.count_nonzero(s != t))
    bytes_total = 8 * stats["expanded_edges"] + 20 * stats["expanded_vertices"] + 4 * V * launches
    avg_launch_ms = float(np.mean(dev_ms)) / launches
    achieved = bytes_total / launches / (avg_launch_ms / 1e3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "visit_kernel<int> (plain two-colour s–t, one query per launch)",
                "bytes_per_launch": int(bytes_total / launches), "avg_launch_us": round(avg_launch_ms * 1e3, 2)}

    # ---- BFS GTEPS ----
    bfs = None
    if args.only in ("all", "bfs"):
        for r in bfs_roots[:2]:
            g.bfs(int(r), levels=False)
        e_tot, ms_tot, lv = 0, 0.0, []
        for r in bfs_roots:
            _, _, st = g.bfs(int(r), levels=False)
            e_tot += st["visited_edges"]
            ms_tot += st["time_ms"]
            lv.append(st["levels"])
        gteps = e_tot / (ms_tot / 1e3) / 1e9
        bfs_bytes = None
        bfs = {"roots": len(bfs_roots), "gteps": round(gteps, 2), "gteps_graph500": round(gteps / 2, 2),
               "avg_ms": round(ms_tot / len(bfs_roots), 3), "avg_levels": float(np.mean(lv))}
        log(f"bfs: {gteps:.1f} GTEPS over {len(bfs_roots)} roots, {ms_tot / len(bfs_roots):.3f} ms avg")

    # ---- CPU baseline: oracle host BFS s–t, one thread, bounded sample ----
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc

        sec1, d1 = orc.time_stcon(hg.offsets, hg.columns, s[:1], t[:1])
        n = int(max(1, min(args.pairs, args.cpu_seconds / max(sec1, 1e-6))))
        sec, dcpu = orc.time_stcon(hg.offsets, hg.columns, s[:n], t[:n])
        assert np.array_equal(dcpu, ref.dist[:n]), "GPU distances differ from the host BFS"
        cpu = {"value": round(n / sec, 3), "unit": "queries/s", "cores": 1, "kind": "port",
               "sample": f"first {n} of the {args.pairs} pairs, host BFS from s then level of t "
                         f"(oracle restatement of BFS-4K Graph::bfs), {sec:.1f}s",
               "cpu": cpu_model(), "nproc": os.cpu_count()}
        log(f"cpu baseline: {n} queries in {sec:.1f}s = {n / sec:.2f} q/s")

    out = {
        "metric": "s–t queries/s + GTEPS, R-MAT s22; batched+learned vs plain, 1 MI355X",
        "value": round(qps, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded R-MAT generator, seeded pairs among degree>0 vertices)",
        "config": {"workload": f"{args.pairs} s–t queries on undirected R-MAT s{args.scale} ef{int(args.edgefactor)}",
                   "scale": args.scale, "edgefactor": args.edgefactor, "V": V, "E_directed": E,
                   "pairs": args.pairs, "path": path, "seed": args.seed},
        "plain_qps": round(plain_qps, 1),
        "bfs": bfs,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
