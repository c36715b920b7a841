"""Time the host and GPU graph builders on one R-MAT graph.

    python -m fastcon.build_bench [--scale 24] [--edgefactor 16] [--no-host]

Prints one line per builder (wall seconds, including the copy of the CSR
back to the host for the GPU builder) and checks that both give the same
arrays."""
import argparse
import time

import numpy as np

import fastcon as fc


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--edgefactor", type=float, default=16.0)
    ap.add_argument("--seed", type=int, default=24)
    ap.add_argument("--no-host", action="store_true")
    args = ap.parse_args()
    fc.gen_rmat_gpu(12)  # first device use pays one-off costs
    t0 = time.perf_counter()
    d = fc.gen_rmat_gpu(args.scale, args.edgefactor, seed=args.seed)
    t1 = time.perf_counter()
    print(f"gpu  R-MAT s{args.scale} ef{args.edgefactor:g}: V={d.V} E={d.E} in {t1 - t0:.3f} s", flush=True)
    if not args.no_host:
        t0 = time.perf_counter()
        h = fc.gen_rmat(args.scale, args.edgefactor, seed=args.seed)
        t1 = time.perf_counter()
        same = h.E == d.E and np.array_equal(h.offsets, d.offsets) and np.array_equal(h.columns, d.columns)
        print(f"host R-MAT s{args.scale} ef{args.edgefactor:g}: V={h.V} E={h.E} in {t1 - t0:.3f} s; same={same}",
              flush=True)
        assert same


if __name__ == "__main__":
    main()
