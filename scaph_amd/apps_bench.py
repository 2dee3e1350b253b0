"""WCC and SSSP measurement, resident and out of core, on R-MAT.

    python -m scaph_amd.apps_bench [--scale 24] [--ef 16] [--budgets 640] [--sources 2]

The graph is generated on the device with a host copy kept; the resident
engine runs WCC and SSSP (weights uniform on [1, wmax]), then for each
-max_memory budget (MiB) a paged engine loads the same CSR and runs them
again.  Prints one JSON line: per run the time, rounds, edges relaxed,
host-to-device bytes, PCIe rate and page hit rate (executions of pages
already resident / all page executions), and whether the paged values equal
the resident ones bit for bit.
"""
import argparse
import json
import statistics
import sys

import numpy as np

from . import BFS


def log(*a):
    print("[apps]", *a, file=sys.stderr, flush=True)


def summary(runs, m):
    ms = statistics.median(r["ms"] for r in runs)
    out = {"ms": ms, "gteps": m / (ms / 1e3) / 1e9, "rounds": [r["depth"] for r in runs],
           "edges_relaxed": [r["m_reached"] for r in runs]}
    if runs[0]["page_execs"]:
        out["bytes_h2d"] = [r["bytes_h2d"] for r in runs]
        out["pcie_gb_s"] = [r["bytes_h2d"] / (r["h2d_ms"] / 1e3) / 1e9 if r["h2d_ms"] else None
                            for r in runs]
        out["page_hit_rate"] = [r["page_execs_cached"] / r["page_execs"] for r in runs]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--ef", type=int, default=16)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--wmax", type=int, default=255)
    ap.add_argument("--budgets", default="640", help="-max_memory values in MiB")
    ap.add_argument("--sources", type=int, default=2)
    ap.add_argument("--repeat", type=int, default=2)
    ap.add_argument("--flags", default="", help="more engine flags for the paged runs")
    args = ap.parse_args()

    res = BFS("-max_memory=262144")
    res.gen_rmat_device(args.scale, args.ef, seed=args.seed, keep_host=True)
    off, ngh = res.host_csr()
    n, m = len(off) - 1, len(ngh)
    wgh = np.random.default_rng(args.seed).integers(1, args.wmax + 1, m, dtype=np.uint32)
    res.set_weights(wgh)
    deg = res.out_degrees()
    rng = np.random.default_rng(args.seed)
    srcs = []
    for v in rng.permutation(np.nonzero(deg > 0)[0]):
        if res.run(int(v))["n_reached"] >= 0.01 * n:
            srcs.append(int(v))
            if len(srcs) == args.sources:
                break
    log("graph", n, m, "sources", srcs)
    out = {"workload": f"R-MAT scale {args.scale} edge factor {args.ef} seed {args.seed}; SSSP "
                       f"weights uniform on [1, {args.wmax}]",
           "vertices": n, "edges": m, "sources": srcs, "paged_flags": args.flags}
    want_wcc = None
    want_sssp = {}
    runs = [res.wcc() for _ in range(args.repeat)]
    want_wcc = res.levels()
    out["resident"] = {"wcc": summary(runs, m)}
    sr = []
    for s in srcs:
        sr.append(res.sssp(s))
        want_sssp[s] = res.levels()
    out["resident"]["sssp"] = summary(sr, m)
    log("resident", out["resident"])
    res.close()
    for b in [int(x) for x in args.budgets.split(",") if x]:
        eng = BFS(f"-max_memory={b}", "-build_pull=false", *args.flags.split())
        eng.load_csr(off, ngh, keep_host=False)
        eng.set_weights(wgh)
        info = eng.info()
        runs = [eng.wcc() for _ in range(args.repeat)]
        exact = bool(np.array_equal(eng.levels(), want_wcc))
        rec = {"pages": info["pages"], "resident": bool(info["resident"]),
               "wcc": summary(runs, m)}
        rec["wcc"]["values_equal_resident"] = exact
        sr = []
        exact = True
        for s in srcs:
            sr.append(eng.sssp(s))
            exact &= bool(np.array_equal(eng.levels(), want_sssp[s]))
        rec["sssp"] = summary(sr, m)
        rec["sssp"]["values_equal_resident"] = exact
        out[f"paged_{b}MiB"] = rec
        log(f"paged {b} MiB", rec)
        eng.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
