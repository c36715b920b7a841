"""Benchmark of the s–t connectivity / BFS hot path on one MI355X.

Workload (one "step"): 1,024 seeded s–t queries on an undirected R-MAT
scale-22 graph (edgefactor 16, a=0.57 b=0.19 c=0.19 d=0.05, int32 CSR),
answered by the batched bit-parallel engine (64-query waves) under the
learned scheduler when trained parameters are present, else under the
heuristic scheduler. `value` is queries per second of device time (HIP
events around the kernel launches; the query and result arrays, 12 KB, are
copied outside the timed region), so inputs are resident in HBM.

Also measured in the same run, on the same graph:
  * the plain one-query-per-launch path on the same pairs (q/s and the
    batched/plain speedup), verdicts and distances compared bit for bit;
  * BFS from 64 seeded roots (GTEPS and the HBM roofline of the visit kernel);
  * the CPU baseline: the oracle's single-thread host BFS s–t (BFS from s,
    level of t) on a bounded sample of the same pairs, also checked.

Prints one JSON line on stdout; progress goes to stderr.
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


def batch_bytes(stats, V, off64):
    """Algorithmic bytes of a batched run: per step and active frontier
    vertex 8 (mask) + o (offset pair) + 8 (visited) + 8 (next mask) bytes, per
    active edge 4 (column) + 8 (neighbour mask) bytes, plus the four 8-byte
    mask words per vertex cleared per wave."""
    o = 16 if off64 else 8
    return (12 * stats["active_edges"] + (24 + o) * stats["active_vertices"]
            + 32 * V * stats["waves"])


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
    ap.add_argument("--policy", choices=["auto", "plain", "heuristic", "learned"], default="auto")
    ap.add_argument("--vw", type=int, default=0, help="virtual-warp width (0: autotuned)")
    ap.add_argument("--tile", type=int, default=0, help="hub chunk in edges (0: autotuned)")
    ap.add_argument("--lanes", type=int, default=0, help="concurrent wave lanes (0: the library default, 3)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="bound on the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-bfs", action="store_true")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the phases directly instead of replaying captured hipGraphs "
                         "(for counter-collection runs, which do not support stream capture)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="run one device-to-device copy of this many MiB and exit: a kernel of known "
                         "traffic that scales hardware byte counters in profiling runs")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration_mib > 0:
        # torch is used here only for the calibration kernel, in a process of
        # its own (the engine is not loaded).
        import torch

        n = args.copy_calibration_mib << 20
        src = torch.ones(n // 4, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {n} bytes read and {n} bytes written by one copy kernel")
        return

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

    policy_name = args.policy
    if policy_name == "auto":
        policy_name = "learned" if fc.load_default_sched_params(g) else "heuristic"
    elif policy_name == "learned" and not fc.load_default_sched_params(g):
        log("no trained scheduler parameters found")
        sys.exit(2)
    policy = {"plain": fc.POLICY_PLAIN, "heuristic": fc.POLICY_HEURISTIC, "learned": fc.POLICY_LEARNED}[policy_name]

    # ---- headline: batched s–t, K timed steps over the pairs ----
    def step():
        return g.stcon_batch(s, t, policy=policy, vw=args.vw, tile=args.tile, use_graph=0 if args.no_graph else 1,
                             lanes=args.lanes)

    ref = step()  # first warm-up step; runs the width/tile autotuner unless both are given
    for _ in range(args.warmup - 1):
        step()
    dev_ms, wall_ms = [], []
    stats = None
    for _ in range(args.steps):
        a = time.perf_counter()
        r = step()
        wall_ms.append((time.perf_counter() - a) * 1e3)
        dev_ms.append(r.stats["time_ms"])
        stats = r.stats
        assert np.array_equal(r.dist, ref.dist)
    ms_step = float(np.mean(dev_ms))
    qps = args.pairs / (ms_step / 1e3)
    log(f"batched ({policy_name}): {ms_step:.3f} ms/step device, {np.mean(wall_ms):.3f} ms wall, {qps:.0f} q/s; "
        f"vw={stats['vw']} tile={stats['tile']} launches={stats['launches']} phases={stats['phases']}")

    # Roofline of the dominant kernel (batch_kernel, every phase of the wave
    # loop is one launch of it): algorithmic bytes per launch over the mean
    # launch interval (device time of the timed region / launches). With
    # concurrent lanes, launches of different lanes overlap, so this interval
    # is shorter than one kernel's own duration.
    launches = int(stats["launches"])
    bytes_run = batch_bytes(stats, V, g.offsets64)
    avg_launch_us = ms_step * 1e3 / launches
    achieved = bytes_run / (ms_step / 1e3) / 1e9
    # Measured HBM traffic (FETCH_SIZE and WRITE_SIZE counter runs, scaled by
    # a copy kernel of known size) is kept per step in profiles/; it applies
    # when this run is the workload it was measured on.
    traffic, traffic_src = None, None
    pmc_path = os.path.join(ROOT, "profiles", "batch_kernel_rmat_s22_pmc.json")
    if os.path.exists(pmc_path):
        pmc = json.load(open(pmc_path))
        m = pmc.get("match", {})
        if (m.get("scale"), m.get("edgefactor"), m.get("pairs"), m.get("policy"), m.get("seed")) == \
                (args.scale, args.edgefactor, args.pairs, policy_name, args.seed):
            traffic = int(pmc["traffic_bytes_per_step"] / launches)
            traffic_src = pmc_path[len(ROOT) + 1:]
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_src,
                "kernel": "fcd::batch_kernel<int>", "bytes_per_launch": int(bytes_run / launches),
                "avg_launch_us": round(avg_launch_us, 2), "launches_per_step": launches,
                "working_launches_per_step": int(stats["phases"]), "lanes": args.lanes or 3}

    # ---- plain one-query-per-launch path on the same pairs ----
    plain = None
    if not args.no_plain:
        g.stcon_many(s[:64], t[:64])
        pr = g.stcon_many(s, t, policy=fc.POLICY_PLAIN)
        assert np.array_equal(pr.dist, ref.dist), "batched and plain distances differ"
        p_ms = pr.stats["time_ms"]
        plain_qps = args.pairs / (p_ms / 1e3)
        nq = int(np.count_nonzero(s != t))
        pst = pr.stats
        pbytes = 8 * pst["expanded_edges"] + 20 * pst["expanded_vertices"] + 4 * V * nq
        plain = {"qps": round(plain_qps, 1), "ms": round(p_ms, 3), "speedup_batched": round(qps / plain_qps, 2),
                 "roofline": {"kernel": "fcd::phase_kernel<int> (one query per launch sequence)",
                              "achieved": round(pbytes / (p_ms / 1e3) / 1e9, 1), "unit": "GB/s",
                              "frac": round(pbytes / (p_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}}
        log(f"plain: {p_ms:.1f} ms = {plain_qps:.0f} q/s; batched speedup {qps / plain_qps:.2f}x")

    # ---- BFS GTEPS ----
    bfs = None
    if not args.no_bfs:
        for r0 in bfs_roots[:2]:
            g.bfs(int(r0), levels=False)
        e_tot, v_tot, ms_tot, lv = 0, 0, 0.0, []
        for r0 in bfs_roots:
            _, _, st = g.bfs(int(r0), levels=False)
            e_tot += st["visited_edges"]
            v_tot += st["visited_vertices"]
            ms_tot += st["time_ms"]
            lv.append(st["levels"])
        gteps = e_tot / (ms_tot / 1e3) / 1e9
        o = 28 if g.offsets64 else 20
        b_bfs = 8 * e_tot + o * v_tot + 4 * V * len(bfs_roots)
        bw = b_bfs / (ms_tot / 1e3) / 1e9
        # Measured HBM bytes per BFS (FETCH_SIZE and WRITE_SIZE counter runs of
        # the phase launches and the un-permutation, scaled by a copy kernel of
        # known size), kept in profiles/ for the workload they were measured on.
        # They fall below the algorithmic bytes because bottom-up levels stop
        # scanning a vertex's list at its first frontier neighbour, while the
        # algorithmic count charges every edge.
        b_traffic, b_src = None, None
        bpmc_path = os.path.join(ROOT, "profiles", "phase_kernel_bfs_rmat_s22_pmc.json")
        if os.path.exists(bpmc_path):
            bp = json.load(open(bpmc_path))
            m = bp.get("match", {})
            if (m.get("scale"), m.get("edgefactor"), m.get("roots"), m.get("seed")) == \
                    (args.scale, args.edgefactor, args.roots, args.seed):
                b_traffic, b_src = int(bp["traffic_bytes_per_bfs"]), bpmc_path[len(ROOT) + 1:]
        bfs = {"roots": len(bfs_roots), "gteps": round(gteps, 2), "gteps_graph500": round(gteps / 2, 2),
               "avg_ms": round(ms_tot / len(bfs_roots), 4), "avg_levels": float(np.mean(lv)),
               "roofline": {"bound": "hbm", "achieved": round(bw, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(bw / HBM_PEAK_GBS, 4), "kernel": "fcd::phase_kernel<int> (BFS)",
                            "bytes_per_bfs": int(b_bfs / len(bfs_roots)), "traffic_per_bfs": b_traffic,
                            "traffic_source": b_src}}
        log(f"bfs: {gteps:.1f} GTEPS over {len(bfs_roots)} roots, {ms_tot / len(bfs_roots):.3f} ms avg, "
            f"{bw:.0f} GB/s algorithmic = {bw / HBM_PEAK_GBS:.1%} of HBM peak")

    # ---- CPU baseline: oracle host BFS s–t, one thread, bounded sample ----
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc

        sec1, _ = orc.time_stcon(hg.offsets, hg.columns, s[:1], t[:1])
        n = int(max(1, min(args.pairs, args.cpu_seconds / max(sec1, 1e-6))))
        sec, dcpu = orc.time_stcon(hg.offsets, hg.columns, s[:n], t[:n])
        assert np.array_equal(dcpu, ref.dist[:n]), "GPU distances differ from the host BFS"
        cpu = {"value": round(n / sec, 3), "unit": "queries/s", "cores": 1, "kind": "port",
               "sample": f"first {n} of the {args.pairs} pairs, host BFS from s then level of t "
                         f"(oracle restatement of the reference host BFS), {sec:.1f}s",
               "cpu": cpu_model(), "nproc": os.cpu_count()}
        log(f"cpu baseline: {n} queries in {sec:.1f}s = {n / sec:.2f} q/s")

    out = {
        "metric": "s–t queries/s + GTEPS, R-MAT s22; batched+learned vs plain, 1 MI355X",
        "value": round(qps, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded R-MAT generator, seeded pairs among degree>0 vertices)",
        "config": {"workload": f"{args.pairs} s–t queries, batched in 64-query waves, on undirected R-MAT "
                               f"s{args.scale} ef{int(args.edgefactor)}",
                   "scale": args.scale, "edgefactor": args.edgefactor, "V": V, "E_directed": E,
                   "pairs": args.pairs, "policy": policy_name, "seed": args.seed,
                   "vw": stats["vw"], "tile": stats["tile"]},
        "wall_ms_per_step": round(float(np.mean(wall_ms)), 4),
        "plain": plain,
        "bfs": bfs,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
