"""Train the batched engine's learned scheduler and report what it gains.

    python -m fastcon.train_sched [--out DIR] [--quick] [--evaluate]

1. Trace: on a seeded family of graphs (undirected R-MAT with the initiator
   a varied around 0.57 and edge factors 8 and 16, Erdős–Rényi with mean
   degrees 6 to 16, 2-D grids with 2% of the edges deleted) the engine runs seeded s–t batches along the heuristic
   schedule and, at every traced step decision, times each of the nine
   actions (side s / t / both; top-down at the engine's width, top-down at
   width 64, bottom-up) from that state to the
   end of the 64-query wave (fc_sched_trace).
   Each state is timed --repeats times (the same decisions each time) and
   the median cost is the label.
2. Train: the 12 -> 32 -> 32 -> 9 network is fitted on the GPU to the
   centred log-costs by the engine's own forward/backward/Adam kernels
   (fc_sched_train). Then the training graphs are traced again along the
   learned schedule itself, on new queries, and the network is refitted on
   both sets of rows (one round of dataset aggregation, so that it also sees
   the states its own choices lead to).
   The final fit's row weighting (every row alike, or each by the time at
   stake, its best cost), its epoch count and the margin by which the
   network must beat the heuristic's action are chosen on validation graphs
   (lowest total regret over a grid of the three).
3. Report, on the training rows and on held-out graphs and queries:
   regret (cost of the chosen action over the best action, per traced step)
   of the network, the heuristic and the plain both-sides schedule; and the
   end-to-end queries/s of the plain, heuristic and learned schedules with
   verdicts and distances compared bit for bit.

Writes sched_params.npy (float32, FC_SCHED_PARAMS values),
sched_report.json and the traced rows (sched_rows.npz: training and
validation features, costs and heuristic actions) into --out.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import fastcon as fc  # noqa: E402


def log(*a):
    print(*a, flush=True)


def rmat(scale, a, seed, ef=16):
    b = (1.0 - a - 0.05) / 2.0
    return fc.gen_rmat(scale, ef, a, b, b, seed=seed)


def family(quick):
    """(name, builder, pairs, trace stride, trace cap) for training and
    held-out evaluation."""
    if quick:
        train = [("rmat16_a0.57", lambda: rmat(16, 0.57, 5), 256, 1, 400),
                 ("grid128", lambda: fc.gen_grid(128, 128, 0.02, 7), 128, 8, 120)]
        valid = [("rmat16_a0.60", lambda: rmat(16, 0.60, 10), 256, 1, 400)]
        held = [("rmat16_a0.50", lambda: rmat(16, 0.50, 9), 256, 1, 400)]
        return train, valid, held
    train = [("rmat20_a0.45", lambda: rmat(20, 0.45, 20), 1024, 1, 800),
             ("rmat20_a0.57", lambda: rmat(20, 0.57, 21), 1024, 1, 800),
             ("rmat20_a0.65", lambda: rmat(20, 0.65, 22), 1024, 1, 800),
             ("rmat21_a0.52", lambda: rmat(21, 0.52, 23), 1024, 1, 800),
             ("rmat22_a0.60", lambda: rmat(22, 0.60, 26), 1024, 1, 800),
             ("rmat23_a0.55", lambda: rmat(23, 0.55, 27), 512, 1, 800),
             ("rmat21_a0.57_ef8", lambda: rmat(21, 0.57, 34, ef=8), 1024, 1, 800),
             ("er2M_d8", lambda: fc.gen_er(1 << 21, 8.0, 24), 1024, 1, 800),
             ("er4M_d6", lambda: fc.gen_er(1 << 22, 6.0, 28), 512, 1, 800),
             ("er2M_d16", lambda: fc.gen_er(1 << 21, 16.0, 35), 1024, 1, 800),
             ("grid512", lambda: fc.gen_grid(512, 512, 0.02, 25), 256, 16, 300),
             ("grid1024", lambda: fc.gen_grid(1024, 1024, 0.02, 29), 128, 32, 200)]
    valid = [("rmat21_a0.57", lambda: rmat(21, 0.57, 41), 1024, 1, 800),
             ("er4M_d8", lambda: fc.gen_er(1 << 22, 8.0, 42), 512, 1, 600)]
    held = [("rmat22_a0.57", lambda: rmat(22, 0.57, 22), 1024, 1, 1200),
            ("rmat21_a0.62", lambda: rmat(21, 0.62, 31), 512, 1, 800),
            ("er8M_d8", lambda: fc.gen_er(1 << 23, 8.0, 32), 512, 1, 600),
            ("grid2048", lambda: fc.gen_grid(2048, 2048, 0.02, 33), 256, 64, 200)]
    return train, valid, held


def trace_graph(name, g, hg, pairs, every, cap, seed, policy=fc.POLICY_HEURISTIC, repeats=1):
    t0 = time.time()
    r = hg.sample_roots(2 * pairs, seed)
    s, t = r[:pairs], r[pairs:]
    ref = g.stcon_batch(s, t, policy=fc.POLICY_PLAIN).dist
    costs = []
    for k in range(repeats):
        feats, cost, heur, wave, dist = g.sched_trace(s, t, cap=cap, every=every, policy=policy)
        assert np.array_equal(dist, ref), f"{name}: traced distances differ from the plain schedule"
        if k == 0:
            f0, h0 = feats, heur
        else:  # the same schedule visits the same states
            assert np.array_equal(feats, f0) and np.array_equal(heur, h0), f"{name}: trace {k} took other steps"
        costs.append(cost)
    cost = np.median(np.stack(costs), axis=0)
    spread = float(np.median(np.ptp(np.stack(costs), axis=0) / cost.clip(min=1e-9))) if repeats > 1 else 0.0
    log(f"  {name}: V={hg.V} E={hg.E} pairs={pairs} rows={len(f0)} x{repeats} "
        f"(median relative spread {spread:.3f}; {time.time() - t0:.1f}s)")
    return s, t, f0, cost, h0


def open_graph(build, seed):
    hg = build()
    g = fc.Graph(hg)
    r = hg.sample_roots(128, seed)
    g.stcon_batch(r[:64], r[64:], policy=fc.POLICY_HEURISTIC)  # autotune width/tile on this graph
    return g, hg


def regret(cost, pick):
    best = cost.min(axis=1)
    chosen = cost[np.arange(len(cost)), pick]
    return float(np.mean(chosen / best - 1.0)), float(np.sum(chosen) / np.sum(best) - 1.0)


def end_to_end(g, s, t, w, margin):
    out = {}
    ref = None
    g.sched_set_params(w)
    for name, pol in (("plain", fc.POLICY_PLAIN), ("heuristic", fc.POLICY_HEURISTIC), ("learned", fc.POLICY_LEARNED)):
        g.stcon_batch(s, t, policy=pol, learn_margin=margin)
        ms = []
        for _ in range(3):
            r = g.stcon_batch(s, t, policy=pol, learn_margin=margin)
            ms.append(r.stats["time_ms"])
            if ref is None:
                ref = r.dist
            assert np.array_equal(r.dist, ref), f"{name} schedule changed an answer"
        out[name] = {"ms": round(float(np.median(ms)), 3), "qps": round(len(s) / (np.median(ms) / 1e3), 1)}
    out["learned_vs_heuristic"] = round(out["heuristic"]["ms"] / out["learned"]["ms"], 3)
    out["learned_vs_plain"] = round(out["plain"]["ms"] / out["learned"]["ms"], 3)
    return out


def evaluate(out, held):
    """Re-measure the shipped network's end-to-end comparison on the held-out
    graphs and queries with the engine as it is now, without tracing or
    training again; the regret rows stay those of the training run."""
    w = np.load(os.path.join(out, "sched_params.npy"), allow_pickle=False)
    path = os.path.join(out, "sched_report.json")
    with open(path) as fh:
        report = json.load(fh)
    margin = report["margin"]
    for i, (name, build, pairs, every, cap) in enumerate(held):
        g, hg = open_graph(build, 250 + i)
        r2 = hg.sample_roots(2 * pairs, 300 + i)
        report["end_to_end"][name] = end_to_end(g, r2[:pairs], r2[pairs:], w, margin)
        log(f"  {name}: end to end {json.dumps(report['end_to_end'][name])}")
        g.close()
    report["end_to_end_note"] = ("re-measured by --evaluate with the engine of the present build; the regret rows "
                                 "come from the training run's traces")
    with open(path, "w") as fh:
        json.dump(report, fh, indent=1)
    log(f"updated {path}")


def pick_actions(pred, heur, margin):
    """The learned policy's choice: the network's best action when it is
    predicted cheaper than the heuristic's by more than margin."""
    best = np.argmin(pred, axis=1)
    idx = np.arange(len(pred))
    return np.where(pred[idx, best] + margin < pred[idx, heur], best, heur)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=HERE)
    ap.add_argument("--quick", action="store_true", help="small graphs (a test-sized run)")
    ap.add_argument("--epochs", type=int, default=3000)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--repeats", type=int, default=3, help="timed rollouts per traced state (median)")
    ap.add_argument("--epochs-grid", default="200,600,2000",
                    help="final-fit epoch counts tried; the one with the lowest validation regret is kept")
    ap.add_argument("--evaluate", action="store_true",
                    help="only re-measure the shipped parameters end to end on the held-out graphs")
    args = ap.parse_args()
    train, valid, held = family(args.quick)
    if args.evaluate:
        evaluate(args.out, held)
        return

    log("trace (training family, heuristic schedule)")
    F, Cst, Hh = [], [], []
    graphs = []
    for i, (name, build, pairs, every, cap) in enumerate(train):
        g, hg = open_graph(build, 50 + i)
        s, t, f, c, h = trace_graph(name, g, hg, pairs, every, cap, 100 + i, repeats=args.repeats)
        F.append(f), Cst.append(c), Hh.append(h)
        graphs.append((name, g, hg, pairs, every, cap))
    g0 = graphs[0][1]
    loss = g0.sched_train(np.concatenate(F), np.concatenate(Cst), epochs=args.epochs, lr=args.lr, seed=args.seed)
    log(f"  first fit: loss {loss:.4f} on {sum(len(f) for f in F)} rows")
    w1 = g0.sched_get_params()

    log("trace (training family, learned schedule)")
    for i, (name, g, hg, pairs, every, cap) in enumerate(graphs):
        g.sched_set_params(w1)
        s, t, f, c, h = trace_graph(name, g, hg, pairs, every, cap, 150 + i, policy=fc.POLICY_LEARNED,
                                    repeats=args.repeats)
        F.append(f), Cst.append(c), Hh.append(h)
    F, Cst, Hh = np.concatenate(F), np.concatenate(Cst), np.concatenate(Hh)

    # validation rows (heuristic schedule) for the choice of epochs and margin
    log("validation graphs")
    VF, VC, VH = [], [], []
    for i, (name, build, pairs, every, cap) in enumerate(valid):
        g, hg = open_graph(build, 350 + i)
        s, t, f, c, h = trace_graph(name, g, hg, pairs, every, cap, 400 + i, repeats=args.repeats)
        VF.append(f), VC.append(c), VH.append(h)
        g.close()
    VF, VC, VH = np.concatenate(VF), np.concatenate(VC), np.concatenate(VH)
    os.makedirs(args.out, exist_ok=True)
    np.savez(os.path.join(args.out, "sched_rows.npz"), features=F, cost=Cst, heuristic=Hh,
             valid_features=VF, valid_cost=VC, valid_heuristic=VH)
    margins = [0.0, 0.02, 0.05, 0.1, 0.15, 0.2, 0.3, 0.5, 1.0]
    weightings = {"uniform": fc.SCHED_WEIGHT_UNIFORM, "time": fc.SCHED_WEIGHT_TIME}
    best = None
    grid = {}
    for wname, wcode in weightings.items():
        for ep in [int(x) for x in args.epochs_grid.split(",") if x.strip()]:
            t0 = time.time()
            loss_ep = g0.sched_train(F, Cst, epochs=ep, lr=args.lr, seed=args.seed, weighting=wcode)
            w_ep = g0.sched_get_params()
            vpred = g0.sched_forward(VF)
            vreg = {m: regret(VC, pick_actions(vpred, VH, m))[1] for m in margins}
            m_ep = min(margins, key=lambda m: (round(vreg[m], 4), -m))
            key_name = f"{wname}/{ep}"
            grid[key_name] = {"loss": round(float(loss_ep), 4), "margin": m_ep,
                              "validation_total_regret_by_margin": {str(m): round(v, 4) for m, v in vreg.items()}}
            log(f"  fit {key_name} epochs on {len(F)} rows: loss {loss_ep:.4f} ({time.time() - t0:.1f}s); validation "
                f"total regret by margin {json.dumps(grid[key_name]['validation_total_regret_by_margin'])}; "
                f"best margin {m_ep}")
            key = (round(vreg[m_ep], 4), -m_ep)
            if best is None or key < best[0]:
                best = (key, wname, ep, m_ep, w_ep, loss_ep, vreg)
    _, weighting, epochs, margin, w, loss, vreg = best
    g0.sched_set_params(w)
    log(f"chosen: {weighting} weighting, {epochs} epochs, margin {margin}; heuristic validation regret "
        f"{regret(VC, VH)[1]:.4f}")

    def report_rows(feats, cost, heur):
        pred = g0.sched_forward(feats)
        rows = {"rows": int(len(feats))}
        for name, p in (("learned", pick_actions(pred, heur, margin)), ("network_argmin", np.argmin(pred, 1)),
                        ("heuristic", heur), ("plain", np.full(len(cost), 2))):
            mr, tr = regret(cost, p)
            rows[name] = {"mean_regret": round(mr, 4), "total_regret": round(tr, 4)}
        rows["best_fixed_action"] = int(np.argmin(cost.sum(axis=0)))
        return rows

    report = {"train": report_rows(F, Cst, Hh), "validation": report_rows(VF, VC, VH), "held_out": {},
              "end_to_end": {}, "loss": float(loss), "weighting": weighting, "epochs": epochs, "lr": args.lr,
              "margin": margin,
              "validation_total_regret_by_margin": {str(m): round(v, 4) for m, v in vreg.items()},
              "epochs_grid": grid, "first_fit_epochs": args.epochs, "repeats": args.repeats}
    log("train-set regret " + json.dumps(report["train"]))
    for name, g, hg, pairs, every, cap in graphs:
        if g is not g0:
            g.close()
    log("held-out graphs")
    for i, (name, build, pairs, every, cap) in enumerate(held):
        g, hg = open_graph(build, 250 + i)
        s, t, f, c, h = trace_graph(name, g, hg, pairs, every, cap, 200 + i, repeats=args.repeats)
        report["held_out"][name] = report_rows(f, c, h)
        r2 = hg.sample_roots(2 * pairs, 300 + i)
        report["end_to_end"][name] = end_to_end(g, r2[:pairs], r2[pairs:], w, margin)
        log(f"  {name}: regret {json.dumps(report['held_out'][name])}")
        log(f"  {name}: end to end {json.dumps(report['end_to_end'][name])}")
        g.close()
    g0.close()

    np.save(os.path.join(args.out, "sched_params.npy"), w.astype(np.float32), allow_pickle=False)
    with open(os.path.join(args.out, "sched_report.json"), "w") as fh:
        json.dump(report, fh, indent=1)
    log(f"wrote {args.out}/sched_params.npy and sched_report.json")


if __name__ == "__main__":
    main()
