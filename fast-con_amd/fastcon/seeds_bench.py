"""Time the helper-seed s–t verdict against the plain visit on a high-diameter
grid.

    python -m fastcon.seeds_bench [--side 4096] [--pairs 32] [--helpers 318] [--no-plain]

The grid has 2% of its edges deleted (seeded); the pairs are at Manhattan
distance 1,000–6,000. Prints the plain visit's and the helper-seed visit's
queries/s (device time) and checks every verdict against the plain visit's,
or with --no-plain against the host graph's connected components."""
import argparse

import numpy as np

import fastcon as fc


def grid_pairs(side, n, lo, hi, seed):
    """n seeded pairs of grid cells whose Manhattan distance is in [lo, hi]."""
    rng = np.random.default_rng(seed)
    s, t = [], []
    while len(s) < n:
        r0, c0 = int(rng.integers(0, side)), int(rng.integers(0, side))
        dr, dc = int(rng.integers(-hi, hi + 1)), int(rng.integers(-hi, hi + 1))
        r1, c1 = r0 + dr, c0 + dc
        if lo <= abs(dr) + abs(dc) <= hi and 0 <= r1 < side and 0 <= c1 < side:
            s.append(r0 * side + c0)
            t.append(r1 * side + c1)
    return np.array(s, np.int32), np.array(t, np.int32)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--side", type=int, default=4096)
    ap.add_argument("--pairs", type=int, default=32)
    ap.add_argument("--helpers", type=int, default=318)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--no-plain", action="store_true", help="skip the plain visit; check against host components")
    args = ap.parse_args()
    hg = fc.gen_grid(args.side, args.side, 0.02, 7)
    s, t = grid_pairs(args.side, args.pairs, 1000, 6000, args.seed)
    with fc.Graph(hg) as g:
        if args.no_plain:
            comp = hg.components()
            want = comp[s] == comp[t]
            g.stcon_seeds(int(s[0]), int(t[0]), helpers=args.helpers, seed=5)  # one-off costs
        else:
            g.stcon_many(s[:2], t[:2])  # first device use pays one-off costs
            plain = g.stcon_many(s, t)
            p_ms = plain.stats["time_ms"]
            print(f"plain visit, {args.side}^2 grid: {len(s)} pairs in {p_ms:.1f} ms = "
                  f"{len(s) / (p_ms / 1e3):.1f} q/s", flush=True)
            want = plain.connected.astype(bool)
        tot, levels = 0.0, 0
        for i in range(len(s)):
            conn, st = g.stcon_seeds(int(s[i]), int(t[i]), helpers=args.helpers, seed=5)
            assert conn == bool(want[i]), f"pair {i}: helper-seed verdict differs"
            tot += st["time_ms"]
            levels += st["levels"]
        print(f"helper seeds ({args.helpers} helpers): {tot:.1f} ms = {len(s) / (tot / 1e3):.1f} q/s, "
              f"{levels / len(s):.0f} levels on average; verdicts agree", flush=True)


if __name__ == "__main__":
    main()
