"""Out-of-core BFS measurement on the power-law digraph: 105M vertices,
~3.7B edges (out-degree Pareto with density ~ x^-1.9, x_min 1.2, capped at
10^6; uniform destinations; seed 3), device page pool capped by
-max_memory, pages streamed from pinned host memory.

    python -m scaph_amd.outofcore_bench [--budgets 15258,8192] [--pools 16e9] [--sources 4]
                                        [--predict off,learned,exact] [--oracle]

--budgets are -max_memory values: the whole device budget, from which the
engine (as src/app.cuh:281-327) first takes the per-vertex and per-page
arrays and gives the rest to the page pool.  --pools instead names the page
pool's own size in bytes (the "16 GB page pool" reading): the -max_memory
that leaves exactly that pool is derived from the vertex and page counts.

Prints one JSON line: per budget the median GTEPS, page hit rate
(executions of already-resident pages / all page executions), host-to-device
bytes and PCIe rates, and parity of the levels with a resident (push-only)
run of the same graph and (--oracle) with the oracle BFS for every source.
--predict runs each pool once per page-activity prediction mode
(-page_predict): off (first free slot), learned (the committed model's
next-level page-activity output) and exact (the levels of a previous run:
the bound for any predictor).
"""
import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

from . import BFS

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PAGE_WORDS = 10 << 20  # the default -pagesize (10 Mi u32 words)


def log(*a):
    print("[c3]", *a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=105_000_000)
    ap.add_argument("--x-min", type=float, default=1.2)
    ap.add_argument("--alpha", type=float, default=1.9)
    ap.add_argument("--cap", type=int, default=1_000_000)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--budgets", default="15258,8192",
                    help="-max_memory values in MiB (15258 MiB = 16e9 bytes)")
    ap.add_argument("--pools", default="",
                    help="page pool sizes in bytes (e.g. 16e9); -max_memory is derived")
    ap.add_argument("--sources", type=int, default=4)
    ap.add_argument("--reps", type=int, default=3,
                    help="paged runs per source, one after the other (the reference's bench "
                         "repeats one source); the median per source is kept")
    ap.add_argument("--threshold", type=int, default=16,
                    help="-threshold of the paged runs in 1/32 pages: 16 compacts pages whose "
                         "active data is at most half a page (the reference's FLAGS_threshold is "
                         "8; app/bfs.cu forces 3)")
    ap.add_argument("--flags", default="", help="more engine flags for the paged runs")
    ap.add_argument("--oracle", action="store_true",
                    help="compare every source's resident levels with the oracle BFS")
    ap.add_argument("--predict", default="off",
                    help="comma-separated -page_predict modes run on every pool")
    ap.add_argument("--schedules", action="store_true",
                    help="also run the resident graph with the pull step: the fixed alpha/beta "
                         "rule against the learned schedule (committed model) on the sources")
    args = ap.parse_args()

    # resident reference run first: builds the graph, picks the sources
    t0 = time.perf_counter()
    res = BFS("-max_memory=262144", "-build_pull=false")
    res.gen_zipf(args.n, args.x_min, args.alpha, args.cap, args.seed)
    info = res.info()
    log("graph", info, f"{time.perf_counter() - t0:.1f}s")
    deg = res.out_degrees()
    rng = np.random.default_rng(args.seed)
    cand = np.nonzero(deg > 0)[0]
    sources, want = [], {}
    for v in rng.permutation(cand)[:50 * args.sources]:
        st = res.run(int(v))
        if st["n_reached"] >= 0.01 * info["n"]:
            sources.append(int(v))
            want[int(v)] = (res.levels(), st)
            if len(sources) == args.sources:
                break
    log("sources", sources, {s: round(want[s][1]["ms"], 3) for s in sources})
    oracle_ok = None
    if args.oracle:
        sys.path.insert(0, ROOT)
        from tests import _oracle as O
        off, ngh = res.host_csr()
        t1 = time.perf_counter()
        oracle_ok = all(bool(np.array_equal(O.bfs_mt(off, ngh, s), want[s][0])) for s in sources)
        log("oracle BFS", oracle_ok, f"{time.perf_counter() - t1:.1f}s")
        del off, ngh
    res_ms = [want[s][1]["ms"] for s in sources]
    res.close()
    schedules = None
    if args.schedules:
        from . import learn as L
        eng = BFS("-max_memory=262144")
        eng.gen_zipf(args.n, args.x_min, args.alpha, args.cap, args.seed)
        L.model_for_graph(eng, L.DEFAULT_MODEL)
        cmp_ = L.compare(eng, sources, repeats=3)
        exact = True
        eng.set_schedule("fixed")
        for s in sources:
            eng.run(s)
            exact &= bool(np.array_equal(eng.levels(), want[s][0]))
        fx = statistics.median(c["fixed_ms"] for c in cmp_)
        ln = statistics.median(c["learned_ms"] for c in cmp_)
        schedules = {"per_source": cmp_, "fixed_ms_median": fx, "learned_ms_median": ln,
                     "learned_speedup_vs_fixed": fx / ln,
                     "learned_speedup_vs_fixed_total":
                         sum(c["fixed_ms"] for c in cmp_) / sum(c["learned_ms"] for c in cmp_),
                     "slowest_source_learned_over_fixed":
                         max(c["learned_ms"] / c["fixed_ms"] for c in cmp_),
                     "push_only_ms_median": statistics.median(res_ms),
                     "fixed_speedup_vs_push_only": statistics.median(res_ms) / fx,
                     "levels_equal_push_only": exact, "model": "scaph_amd/schedule_mlp.bin"}
        log("schedules", schedules)
        eng.close()

    out = {"workload": f"power-law digraph n={args.n} x_min={args.x_min} alpha={args.alpha} "
                       f"cap={args.cap} seed={args.seed}",
           "vertices": info["n"], "edges": info["m"],
           "resident_gteps_median": statistics.median(
               want[s][1]["m_reached"] / (want[s][1]["ms"] / 1e3) / 1e9 for s in sources),
           "resident_ms_median": statistics.median(res_ms),
           "levels_resident_vs_oracle": oracle_ok,
           "oracle_sources": len(sources) if args.oracle else 0,
           "schedules": schedules, "budgets": []}
    slot = PAGE_WORDS * 4
    budgets = [(int(x), None) for x in args.budgets.split(",") if x]
    for x in [float(x) for x in args.pools.split(",") if x]:
        # plan_residency: pool = (budget - 9 n - 24 pages) / slot, whole slots
        need = int(x) // slot * slot + 9 * info["n"] + 24 * info["pages"]
        budgets.append(((need + (1 << 20) - 1) >> 20, int(x)))
    host = None
    for b, pool_bytes in budgets:
        for mode in [m for m in args.predict.split(",") if m]:
            paged_flags = [f"-max_memory={b}", "-build_pull=false", f"-threshold={args.threshold}",
                           f"-page_predict={mode}"]
            paged_flags += args.flags.split()
            eng = BFS(*paged_flags)
            t0 = time.perf_counter()
            if host is None:
                eng.gen_zipf(args.n, args.x_min, args.alpha, args.cap, args.seed)
                host = eng.host_csr()
            else:
                eng.load_csr(*host)
            inf = eng.info()
            log("budget", b, mode, inf, f"{time.perf_counter() - t0:.1f}s")
            if mode == "learned":
                from . import learn as L
                L.model_for_graph(eng, L.DEFAULT_MODEL)
            runs, exact = [], True
            if mode == "exact":
                eng.set_page_levels(want[sources[0]][0])
            eng.run(sources[0])  # warm: first touches of the pinned pages
            per_src = []
            for s in sources:
                if mode == "exact":
                    eng.set_page_levels(want[s][0])
                rep = []
                for _ in range(args.reps):
                    st = eng.run(s)
                    exact &= bool(np.array_equal(eng.levels(), want[s][0]))
                    runs.append(st)
                    rep.append(st["ms"])
                    log(" ", s, {k: st[k] for k in ("ms", "page_execs", "page_execs_cached",
                                                     "pages_full", "pages_compacted",
                                                     "pages_prefetched", "pages_direct",
                                                     "bytes_h2d", "h2d_ms")})
                per_src.append(statistics.median(rep))
            ms = statistics.median(per_src)
            execs = sum(r["page_execs"] for r in runs)
            cached = sum(r["page_execs_cached"] for r in runs)
            h2d = sum(r["bytes_h2d"] for r in runs)
            h2d_ms = sum(r["h2d_ms"] for r in runs)
            tot_ms = sum(r["ms"] for r in runs)
            slots = eng.pool()["slots"]
            out["budgets"].append({
                "max_memory_mib": b, "pool_bytes_requested": pool_bytes, "page_predict": mode,
                "pool_slots": slots, "pool_bytes": slots * slot,
                "flags": paged_flags, "resident": inf["resident"],
                "pages": inf["pages"],
                "gteps_median": statistics.median(r["m_reached"] / (r["ms"] / 1e3) / 1e9
                                                  for r in runs),
                "t_bfs_ms_median": ms,
                "t_bfs_ms_per_source": per_src,
                "t_bfs_ms": [r["ms"] for r in runs],
                "page_hit_rate": cached / execs if execs else None,
                "page_execs": execs, "pages_full_copies": sum(r["pages_full"] for r in runs),
                "pages_compacted": sum(r["pages_compacted"] for r in runs),
                "pages_copied_ahead": sum(r["pages_prefetched"] for r in runs),
                "pages_direct": sum(r["pages_direct"] for r in runs),
                "h2d_bytes_per_bfs": h2d / len(runs),
                "pcie_gbps_over_bfs": h2d / (tot_ms / 1e3) / 1e9 if tot_ms else None,
                "pcie_gbps_during_copies": h2d / (h2d_ms / 1e3) / 1e9 if h2d_ms else None,
                "levels_equal_resident": exact})
            eng.close()
    # speed of each prediction mode against off on the same pool
    base = {(x["max_memory_mib"]): x["t_bfs_ms_median"] for x in out["budgets"]
            if x["page_predict"] == "off"}
    for x in out["budgets"]:
        if x["max_memory_mib"] in base:
            x["speedup_vs_off"] = base[x["max_memory_mib"]] / x["t_bfs_ms_median"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
