"""Extract the known-answer rows of the reference's published ST-CON run on
R-MAT scale 16 (bernaschi/ST_CONN_ONE_ATOMIC/current-st/st-conn/data/
st_stats_20200418_140644, header lines 1-8, rows 9-108) into a small JSON
fixture: seeds, and per query (run, RED-V, BLUE-V, DEPTH, MN-FOUND).
Reads the data file as text; runs nothing from the reference."""
import json
import os
import sys

SRC = "/root/reference/bernaschi/ST_CONN_ONE_ATOMIC/current-st/st-conn/data/st_stats_20200418_140644"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "st_stats_s16.json")


def main():
    lines = open(SRC).read().splitlines()
    hdr = dict(l.split(None, 1) for l in lines[:7] if " " in l)
    rows = []
    for l in lines[8:108]:
        run, red, blue, depth, mn = l.split()[:5]
        rows.append([int(run), int(red), int(blue), int(depth), int(mn)])
    fx = {
        "source": "st-conn/data/st_stats_20200418_140644:1-108",
        "scale": int(hdr["SCALE"]),
        "edgefactor": float(hdr["Edgefactor"]),
        "seed1": int(hdr["Seed1"]),
        "seed2": int(hdr["Seed2"]),
        "root_seeds": [23, 24],
        "columns": ["run", "red_v", "blue_v", "depth", "mn_found"],
        "rows": rows,
        "max_level": int(lines[108].split()[1]),
    }
    json.dump(fx, open(OUT, "w"), indent=0)
    print(f"wrote {OUT}: {len(rows)} rows")


if __name__ == "__main__":
    sys.exit(main())
