"""A/B comparison of two engine flag sets on one device-built R-MAT graph
and the same sources (min of `reps` runs per source), e.g.

    python -m scaph_amd.ab --a="-pull_head=false" --b="-pull_head=true"

Prints one JSON line with both medians, the median per-source ratio B/A and
whether the two engines' final levels agree.
"""
import argparse
import json
import statistics

import numpy as np

from . import BFS


def pick_sources(eng, count, seed, min_frac=0.01):
    deg = eng.out_degrees()
    n = len(deg)
    cand = np.nonzero(deg > 0)[0]
    rng = np.random.default_rng(seed)
    out = []
    for v in rng.permutation(cand)[:20 * count]:
        if eng.run(int(v))["n_reached"] >= min_frac * n:
            out.append(int(v))
            if len(out) == count:
                break
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--a", default="", help="engine flags of A (space separated)")
    ap.add_argument("--b", default="", help="engine flags of B (space separated)")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--scale", type=int, default=26)
    ap.add_argument("--ef", type=int, default=32)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--sources", type=int, default=32)
    args = ap.parse_args()
    fa = args.a.split()
    fb = args.b.split()
    common = ["-max_memory=262144", "-pagesize=10"]
    ea = BFS(*common, *fa)
    ea.gen_rmat_device(args.scale, args.ef, 0.57, 0.19, 0.19, seed=args.seed)
    eb = BFS(*common, *fb)
    eb.gen_rmat_device(args.scale, args.ef, 0.57, 0.19, 0.19, seed=args.seed)
    srcs = pick_sources(ea, args.sources, args.seed)
    ta, tb = [], []
    equal = True
    for s in srcs:
        ta.append(min(ea.run(s)["ms"] for _ in range(args.reps)))
        tb.append(min(eb.run(s)["ms"] for _ in range(args.reps)))
        equal &= bool(np.array_equal(ea.levels(), eb.levels()))
    ratios = [b / a for a, b in zip(ta, tb)]
    print(json.dumps({"A": fa, "B": fb, "sources": len(srcs), "median_A_ms": statistics.median(ta),
                      "median_B_ms": statistics.median(tb),
                      "median_ratio_B_over_A": statistics.median(ratios),
                      "total_ratio_B_over_A": sum(tb) / sum(ta),
                      "ratio_quartiles": statistics.quantiles(ratios, n=4) if len(ratios) > 1
                      else ratios,
                      "levels_equal": equal}), flush=True)


if __name__ == "__main__":
    main()
