"""Benchmark of the s–t connectivity / BFS hot path on one MI355X.

Workload (one "step"): 1,024 seeded s–t queries on an undirected R-MAT
scale-22 graph (edgefactor 16, a=0.57 b=0.19 c=0.19 d=0.05, int32 CSR),
answered by the batched bit-parallel engine (64-query waves) under the
heuristic schedule; the learned schedule (shipped parameters, validated
margin) is timed beside it on the same queries, and at that margin follows
the heuristic's decisions. `value` is queries per second of device time (HIP
events around the kernel launches; the query and result arrays, 12 KB, are
copied outside the timed region), so inputs are resident in HBM.

Also measured in the same run, on the same graph:
  * the plain one-query-per-launch path on the same pairs (q/s and the
    batched/plain speedup), verdicts and distances compared bit for bit;
  * BFS from 64 seeded roots (GTEPS and the HBM roofline of the visit kernel);
  * the one published per-query figure of the path (ST-CON, R-MAT s16,
    0.579 ms per query) beside this engine's plain visit on an R-MAT s16;
  * the CPU baseline: the reference's single-thread host BFS (restated in
    the oracle, built -O3 -march=native where this runs): s–t queries/s on a
    bounded sample of the same pairs (checked against the GPU), BFS GTEPS,
    and s16 ms per query; one warm-up, then the median of three repetitions.

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
    ap.add_argument("--policy", choices=["plain", "heuristic", "learned"], default="heuristic",
                    help="schedule of the headline batched run; the learned schedule is also timed beside it")
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
    have_learned = fc.load_default_sched_params(g)
    if policy_name == "learned" and not have_learned:
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

    # The learned schedule on the same queries (shipped parameters and the
    # margin chosen with them on validation graphs). At that margin the
    # network takes the heuristic's decisions on held-out graphs, so the
    # headline runs the heuristic and the learned schedule is reported here.
    learned = None
    if have_learned and policy_name != "learned":
        lms = []
        for i in range(args.steps + 1):
            rl = g.stcon_batch(s, t, policy=fc.POLICY_LEARNED, vw=stats["vw"], tile=stats["tile"],
                               use_graph=0 if args.no_graph else 1, lanes=args.lanes)
            assert np.array_equal(rl.dist, ref.dist)
            if i:
                lms.append(rl.stats["time_ms"])
        l_ms = float(np.mean(lms))
        learned = {"ms": round(l_ms, 4), "qps": round(args.pairs / (l_ms / 1e3), 1),
                   "vs_headline": round(ms_step / l_ms, 3), "margin": g.learn_margin,
                   "note": "shipped parameters at the validated margin: the network follows the heuristic's "
                           "decisions on held-out graphs, so this is the heuristic plus inference cost"}
        log(f"batched (learned, margin {g.learn_margin}): {l_ms:.3f} ms/step, {learned['qps']:.0f} q/s, "
            f"{learned['vs_headline']:.3f}x the {policy_name} schedule")

    # Roofline of the dominant kernel (batch_kernel: every phase of the wave
    # loop is one launch of it). Concurrent lanes overlap their launches, and
    # a profiler serialises them, so the kernel's own average launch duration
    # is measured on one lane: the same queries, timed with HIP events on the
    # lane's stream over the timed steps (device time / launches). The
    # algorithmic bytes per launch are those of that run.
    one = []
    for _ in range(args.steps):
        r1 = g.stcon_batch(s, t, policy=policy, vw=stats["vw"], tile=stats["tile"],
                           use_graph=0 if args.no_graph else 1, lanes=1)
        assert np.array_equal(r1.dist, ref.dist)
        one.append(r1.stats)
    one_ms = float(np.mean([x["time_ms"] for x in one]))
    one_launches = float(np.mean([x["launches"] for x in one]))
    one_bytes = float(np.mean([batch_bytes(x, V, g.offsets64) for x in one]))
    avg_launch_us = one_ms * 1e3 / one_launches
    achieved = one_bytes / one_launches / (avg_launch_us * 1e-6) / 1e9
    log(f"batched, one lane: {one_ms:.3f} ms/step, {one_launches:.0f} launches, {avg_launch_us:.2f} us per launch, "
        f"{achieved:.0f} GB/s algorithmic")
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
            traffic = int(pmc["traffic_bytes_per_step"] / one_launches)
            traffic_src = pmc_path[len(ROOT) + 1:]
    launches = int(stats["launches"])
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_src,
                "kernel": "fcd::batch_kernel<int>", "bytes_per_launch": int(one_bytes / one_launches),
                "avg_launch_us": round(avg_launch_us, 2), "launches_per_step": round(one_launches, 1),
                "measured_on": "one lane (same queries; concurrent lanes overlap launches)",
                "headline_lanes": args.lanes or 3, "headline_launches_per_step": launches,
                "headline_working_launches_per_step": int(stats["phases"]),
                "headline_algorithmic_gbs": round(batch_bytes(stats, V, g.offsets64) / (ms_step / 1e3) / 1e9, 1)}

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

    # ---- s16: the only published per-query figure of the path ----
    # The reference's ST-CON run on R-MAT scale 16 (edgefactor 16) reports a
    # mean of 0.579 ms per s–t query (st-conn/data/st_stats_20200418_140644:110).
    # The same kind of graph from this engine's generator (seed 16) and its
    # 64 seeded pairs (the CPU configuration), plain visit, one query per
    # launch sequence.
    hg16 = fc.gen_rmat(16, 16, 0.57, 0.19, 0.19, seed=16)
    r16 = hg16.sample_roots(128, 64)
    s16, t16 = r16[:64], r16[64:]
    with fc.Graph(hg16) as g16:
        g16.stcon_many(s16[:4], t16[:4])
        p16 = [g16.stcon_many(s16, t16) for _ in range(3)]
    for x in p16[1:]:
        assert np.array_equal(x.dist, p16[0].dist)
    s16_gpu_ms = float(np.median([x.stats["time_ms"] for x in p16])) / len(s16)
    s16_info = {"gpu_plain_ms_per_query": round(s16_gpu_ms, 4),
                "reference_ms_per_query": 0.579,
                "reference_source": "bernaschi/ST_CONN_ONE_ATOMIC/current-st/st-conn/data/st_stats_20200418_140644:110 "
                                    "(Graph500 R-MAT s16 ef16, 100 queries, GPU not recorded)",
                "workload": "this engine's R-MAT s16 ef16 (seed 16), 64 seeded pairs, median of 3 runs"}
    log(f"s16 plain: {s16_gpu_ms:.4f} ms per query (reference's published mean: 0.579 ms)")

    # ---- CPU baseline: the reference's host BFS, one thread ----
    # oracle.c compiled here with -O3 -march=native; one warm-up query, then
    # the median of three timed repetitions of a bounded sample of the same
    # pairs (BFS from s, level of t: the reference's own host path), checked
    # against the GPU's distances; BFS GTEPS from the first BFS roots the same
    # way; and the s16 pairs above.
    cpu = None
    if not args.no_cpu:
        import shutil
        import tempfile

        from oracle import oracle as orc

        work = tempfile.mkdtemp(prefix="fastcon_cpu_")
        lib, build = orc.native_build(work)
        if lib is None:
            build = f"prebuilt -O3 oracle ({build})"
        sec1, _ = orc.time_stcon(hg.offsets, hg.columns, s[:1], t[:1], lib=lib)  # warm-up
        n = int(max(1, min(args.pairs, args.cpu_seconds / 3 / max(sec1, 1e-6))))
        reps = [orc.time_stcon(hg.offsets, hg.columns, s[:n], t[:n], lib=lib) for _ in range(3)]
        for _, dcpu in reps:
            assert np.array_equal(dcpu, ref.dist[:n]), "GPU distances differ from the host BFS"
        sec = float(np.median([x[0] for x in reps]))
        nb = 4
        breps = [orc.time_bfs(hg.offsets, hg.columns, bfs_roots[:nb], lib=lib) for _ in range(3)]
        bsec = float(np.median([x[0] for x in breps]))
        cpu_gteps = breps[0][1] / bsec / 1e9
        c16 = [orc.time_stcon(hg16.offsets, hg16.columns, s16, t16, lib=lib) for _ in range(3)]
        assert np.array_equal(c16[0][1], p16[0].dist), "GPU s16 distances differ from the host BFS"
        s16_info["cpu_ms_per_query"] = round(float(np.median([x[0] for x in c16])) * 1e3 / len(s16), 4)
        shutil.rmtree(work, ignore_errors=True)
        cpu = {"value": round(n / sec, 3), "unit": "queries/s", "cores": 1, "kind": "port",
               "sample": f"first {n} of the {args.pairs} pairs, host BFS from s then level of t (oracle "
                         f"restatement of the reference host BFS), median of 3 repetitions of {sec:.1f}s "
                         f"after one warm-up query",
               "build": build, "bfs_gteps": round(cpu_gteps, 4),
               "bfs_sample": f"{nb} of the BFS roots, median of 3 repetitions of {bsec:.2f}s",
               "cpu": cpu_model(), "nproc": os.cpu_count()}
        log(f"cpu baseline ({build}): {n} queries, median {sec:.2f}s = {n / sec:.2f} q/s; "
            f"BFS {cpu_gteps:.3f} GTEPS; s16 {s16_info['cpu_ms_per_query']:.3f} ms per query")

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
        "learned": learned,
        "bfs": bfs,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "s16": s16_info,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
