"""Time the plain one-query-per-launch s–t visit (fc_stcon_many).

    python -m fastcon.plain_bench [--graph grid|rmat] [--pairs N]

grid: the 4096 x 4096 grid with 2% of its edges deleted (seed 7) and pairs at
Manhattan distance 1,000–6,000 (seed 11); rmat: undirected R-MAT scale 22,
edgefactor 16 (seed 22) with pairs drawn by fc_sample_roots (seed 23), as in
bench.py. One warm-up call on two pairs runs first. Prints queries/s (HIP
events), working launches per query and the device time per launch. The
distances are compared with the batched engine's; their parity with the
host BFS is covered by the tests."""
import argparse

import numpy as np

import fastcon as fc
from fastcon.seeds_bench import grid_pairs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--graph", choices=["grid", "rmat"], default="grid")
    ap.add_argument("--pairs", type=int, default=32)
    ap.add_argument("--no-batch", action="store_true", help="skip the comparison with the batched engine")
    args = ap.parse_args()
    if args.graph == "grid":
        hg = fc.gen_grid(4096, 4096, 0.02, 7)
        s, t = grid_pairs(4096, args.pairs, 1000, 6000, 11)
    else:
        hg = fc.gen_rmat(22, 16, 0.57, 0.19, 0.19, seed=22)
        r = hg.sample_roots(64 + 2 * args.pairs, 23)
        s, t = r[64:64 + args.pairs], r[64 + args.pairs:]
    print(f"{args.graph}: V={hg.V} E={hg.E}, {len(s)} pairs", flush=True)
    with fc.Graph(hg) as g:
        g.stcon_many(s[:2], t[:2])
        res = g.stcon_many(s, t)
        st = res.stats
        ms = st["time_ms"]
        print(f"plain visit: {len(s)} pairs in {ms:.2f} ms = {len(s) / (ms / 1e3):.1f} q/s; "
              f"{st['launches'] / len(s):.1f} working launches and {st['levels'] / len(s):.1f} steps per query, "
              f"{ms * 1e3 / max(st['launches'], 1):.2f} us of device time per working launch", flush=True)
        if args.no_batch:
            return
        bres = g.stcon_batch(s, t)
        assert np.array_equal(res.dist, bres.dist), "plain and batched distances differ"
        print("distances agree with the batched engine", flush=True)


if __name__ == "__main__":
    main()
