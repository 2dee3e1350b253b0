"""Memory-side traffic of the BFS kernels, from rocprofv3 counters.

Two steps:

  rocprofv3 --pmc FETCH_SIZE -d <dir>/fetch -o fetch --output-format csv -- \\
      python3 -m scaph_amd.traffic run --records <dir>/records.json
  rocprofv3 --pmc WRITE_SIZE -d <dir>/write -o write --output-format csv -- \\
      python3 -m scaph_amd.traffic run --records <dir>/records.json
  python3 -m scaph_amd.traffic summarize --dir <dir> --out profiles/<name>.json

and, from a kernel-trace run into <dir>/kt, a text summary:

  python3 -m scaph_amd.traffic stats --dir <dir> --out profiles/<name>.txt \\
      --title "..." --program "<the profiled command>"

`run` first reads and writes a buffer of known size with the library's own
kernels (scaph_checksum_u32 reads 4n bytes, scaph_fill_u32 writes 4n bytes),
then runs one workload, writing every run's per-level records:
  --workload rmat   BFS on R-MAT scale 26, edge factor 32, seed 2 (the
                    benchmark graph), giant-component sources;
  --workload zipf   BFS on the resident power-law digraph (105 M vertices,
                    3.7 B edges, seed 3), push only;
  --workload msbfs  256-query multi-source BFS on R-MAT scale 24, edge
                    factor 16, seed 3: the autotuner's sweep, then four runs
                    of its best per-level launch sequence without graph
                    replay.
`summarize --kernel K` scales FETCH_SIZE and WRITE_SIZE (KiB per dispatch)
by what they read for those two kernels, maps K's dispatches of the timed
runs onto their levels, and writes bytes per launch beside the algorithmic
bytes of the same launches (models in alg_bytes), stamped with the hash of
the kernel sources (bench.py uses a figure only for that build).
"""
import argparse
import csv
import ctypes
import json
import os
import re
import statistics
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CALIB_WORDS = 256 << 20  # 1 GiB


def _calibrate():
    from . import _lib
    hip = _lib.hip()
    buf, out = ctypes.c_void_p(), ctypes.c_void_p()
    _lib.check(hip.scaph_dev_alloc(ctypes.byref(buf), CALIB_WORDS * 4))
    _lib.check(hip.scaph_dev_alloc(ctypes.byref(out), 64))
    try:
        _lib.check(hip.scaph_fill_u32(buf, 1, CALIB_WORDS, None))
        _lib.check(hip.scaph_checksum_u32(buf, CALIB_WORDS, out, None))
        _lib.check(hip.scaph_device_sync())
        s = ctypes.c_uint64()
        _lib.check(hip.scaph_copy_d2h_async(ctypes.byref(s), out, 8, None))
        _lib.check(hip.scaph_device_sync())
        if s.value != CALIB_WORDS:
            raise RuntimeError(f"calibration checksum {s.value} != {CALIB_WORDS}")
    finally:
        hip.scaph_dev_free(buf)
        hip.scaph_dev_free(out)


def run(args):
    sys.path.insert(0, ROOT)
    import bench
    import scaph_amd
    _calibrate()
    if args.workload == "msbfs":
        flags = ["-max_memory=262144", "-pagesize=10"]
        e = scaph_amd.BFS(*flags)
        e.gen_rmat_device(24, 16, 0.57, 0.19, 0.19, seed=3)
        info = e.info()
        srcs, _ = bench.pick_sources(e, info["n"], 256, seed=3, min_reach_frac=0.01)
        if args.ms_tune:
            tune = json.loads(args.ms_tune)
        else:
            # the autotuner's best per-level launch sequence (its launches
            # come first; summarize takes the last dispatches of the kernel)
            _, _, log_ = e.msbfs_autotune(np.array(srcs, np.uint32))
            tune = min((t for t in log_ if not t[0].get("persistent")), key=lambda t: t[1])[0]
            tune["use_graph"] = 0
        runs = []
        for _ in range(4):  # one warm-up, three timed
            st = e.msbfs(srcs, tune)
            runs.append({"ms": st["ms"], "depth": st["depth"], "levels": e.msbfs_counters()})
    else:
        if args.workload == "zipf":
            flags = ["-max_memory=262144", "-build_pull=false"] + args.flag
            e = scaph_amd.BFS(*flags)
            e.gen_zipf(105_000_000, 1.2, 1.9, 1_000_000, 3)
            nsrc = min(args.sources, 4)
        else:
            flags = ["-max_memory=262144", "-pagesize=10"] + args.flag
            e = scaph_amd.BFS(*flags)
            e.gen_rmat_device(args.scale, args.ef, 0.57, 0.19, 0.19, seed=args.seed)
            nsrc = args.sources
        info = e.info()
        srcs, _ = bench.pick_sources(e, info["n"], nsrc,
                                     seed=3 if args.workload == "zipf" else args.seed,
                                     min_reach_frac=0.01)
        runs = []
        for s in srcs[:1] + srcs:  # a warm-up run, then the timed ones
            st = e.run(s)
            runs.append({"src": s, "ms": st["ms"], "n_reached": st["n_reached"],
                         "m_reached": st["m_reached"], "levels": e.records()})
    e.close()
    with open(args.records, "w") as f:
        json.dump({"workload": args.workload, "n": info["n"], "m": info["m"], "flags": flags,
                   "ms_tune": tune if args.workload == "msbfs" else None,
                   "timed_from": 1, "runs": runs}, f)
    print(json.dumps({"runs": len(runs)}), flush=True)


def _kernel(name):
    name = name.replace("(anonymous namespace)::", "")
    if name.startswith("void "):
        name = name[5:]
    return re.split(r"[(<]", name)[0]


def _dispatches(path, counter):
    out = []
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter:
                continue
            out.append({"id": int(r["Dispatch_Id"]), "kernel": _kernel(r["Kernel_Name"]),
                        "kib": float(r["Counter_Value"]),
                        "us": (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3})
    out.sort(key=lambda d: d["id"])
    return out


def alg_bytes(kernel, rec, n, nxt=None):
    """Algorithmic bytes of one launch.
    k_pull (one pull level): 12 B Node per examined vertex, 4 B per scanned
      in-edge, 4 B level write per vertex found, and the visited read,
      visited write and next-frontier write (3 n/8), as bench.py counts them;
    k_expand (one push level past the one-launch size): 12 B Node + 4 B
      queue read per frontier vertex, 4 B per scanned out-edge, 4 B level
      + 4 B queue write per vertex reached;
    k_ms_pull (one multi-source pull level): the 32-B seen read and 32-B
      frontier write of every vertex and the next membership bitmap (n/8),
      plus per newly reached vertex its 32-B seen write and 4-B degree."""
    if kernel == "k_pull":
        return 12 * rec["examined"] + 4 * rec["scanned"] + 4 * rec["n_next"] + 3 * (n / 8)
    if kernel == "k_expand":
        return 16 * rec["n_f"] + 4 * rec["scanned"] + 8 * rec["n_next"]
    return 64 * n + n / 8 + 36 * nxt["n_active"]


def _launches(rec, kernel):
    """The timed levels that launch `kernel`, in launch order."""
    out = []
    timed = rec["runs"][rec["timed_from"]:]
    if kernel == "k_ms_pull":
        # one launch per level (it returns at once on push levels)
        for r in timed:
            lv = r["levels"]
            for L in range(r["depth"]):
                out.append((lv[L], lv[L + 1] if L + 1 < len(lv) else {"n_active": 0}))
        return out
    for r in timed:
        for lv in r["levels"]:
            if kernel == "k_pull" and lv["dir"] == 1 and not lv.get("sparse"):
                out.append((lv, None))
            elif (kernel == "k_expand" and lv["dir"] == 0 and
                  (lv["m_f"] > 8192 or lv["n_f"] > 65536)):
                out.append((lv, None))
    return out


def summarize(args):
    sys.path.insert(0, ROOT)
    import bench
    d = args.dir
    fetch = _dispatches(os.path.join(d, "fetch", "fetch_counter_collection.csv"), "FETCH_SIZE")
    write = _dispatches(os.path.join(d, "write", "write_counter_collection.csv"), "WRITE_SIZE")
    with open(os.path.join(d, "records.json")) as f:
        rec = json.load(f)
    calib_bytes = CALIB_WORDS * 4
    rd = [x for x in fetch if x["kernel"] == "k_checksum_u32"]
    wr = [x for x in write if x["kernel"] == "k_fill_u32"]
    if not rd or not wr:
        raise SystemExit("calibration dispatches missing")
    f_scale = calib_bytes / (rd[0]["kib"] * 1024)
    w_scale = calib_bytes / (max(wr, key=lambda x: x["kib"])["kib"] * 1024)
    kernel = args.kernel
    launches = _launches(rec, kernel)
    fk = [x for x in fetch if x["kernel"] == kernel]
    wk = [x for x in write if x["kernel"] == kernel]
    if len(fk) != len(wk) or len(fk) < len(launches):
        raise SystemExit(f"{kernel} dispatches: {len(fk)} fetch, {len(wk)} write, "
                         f"{len(launches)} timed launches")
    fk, wk = fk[-len(launches):], wk[-len(launches):]
    per = []
    for (lv, nxt), a, b in zip(launches, fk, wk):
        fb = a["kib"] * 1024 * f_scale
        wb = b["kib"] * 1024 * w_scale
        d = {k: lv[k] for k in ("level", "n_f", "examined", "scanned", "n_next", "n_active")
             if k in lv}
        d.update({"fetch_bytes": fb, "write_bytes": wb,
                  "alg_bytes": alg_bytes(kernel, lv, rec["n"], nxt), "us_profiled": a["us"]})
        per.append(d)
    if kernel == "k_ms_pull":
        # launches that returned at once (push levels) are not pull launches
        per = [p for p in per if p["us_profiled"] > 20.0]
    k = len(per)
    traffic = sum(p["fetch_bytes"] + p["write_bytes"] for p in per) / k
    alg = sum(p["alg_bytes"] for p in per) / k
    avg_us = statistics.mean(p["us_profiled"] for p in per)
    command = args.command or (
        "rocprofv3 --pmc FETCH_SIZE --output-format csv -d <dir>/fetch -o fetch -- python3 -m "
        f"scaph_amd.traffic run --workload {rec.get('workload', 'rmat')} --records "
        "<dir>/records.json; rocprofv3 --pmc WRITE_SIZE --output-format csv -d <dir>/write -o "
        "write -- (the same run); python3 -m scaph_amd.traffic summarize --dir <dir> --kernel "
        f"{kernel} --out {args.out}")
    out = {"command": command, "workload": rec.get("workload", "rmat"), "kernel": kernel,
           "launches": k, "kernel_source_hash": bench.kernel_source_hash(),
           "calibration": {"read_kernel": "k_checksum_u32", "write_kernel": "k_fill_u32",
                           "bytes": calib_bytes, "fetch_raw_bytes": rd[0]["kib"] * 1024,
                           "write_raw_bytes": max(wr, key=lambda x: x["kib"])["kib"] * 1024,
                           "fetch_scale": f_scale, "write_scale": w_scale},
           "fetch_bytes_per_launch": sum(p["fetch_bytes"] for p in per) / k,
           "write_bytes_per_launch": sum(p["write_bytes"] for p in per) / k,
           "traffic_bytes_per_launch": traffic, "alg_bytes_per_launch": alg,
           "traffic_over_alg": traffic / alg if alg else None,
           "avg_launch_us_profiled": avg_us,
           "alg_gbps": alg / (avg_us * 1e-6) / 1e9,
           "alg_frac_of_8tbps": alg / (avg_us * 1e-6) / 8e12,
           "traffic_gbps": traffic / (avg_us * 1e-6) / 1e9,
           "per_launch": per}
    # every kernel of the timed region, for the other rows
    kern = {}
    for x, y in zip(fetch, write):
        if x["kernel"] != y["kernel"]:
            continue
        t = kern.setdefault(x["kernel"], {"dispatches": 0, "fetch_bytes": 0.0,
                                          "write_bytes": 0.0, "us": 0.0})
        t["dispatches"] += 1
        t["fetch_bytes"] += x["kib"] * 1024 * f_scale
        t["write_bytes"] += y["kib"] * 1024 * w_scale
        t["us"] += x["us"]
    out["kernels_whole_process"] = dict(sorted(kern.items(), key=lambda kv: -kv[1]["us"])[:16])
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps({k_: out[k_] for k_ in ("kernel", "launches", "fetch_bytes_per_launch",
                                             "write_bytes_per_launch", "alg_bytes_per_launch",
                                             "traffic_over_alg", "avg_launch_us_profiled",
                                             "alg_frac_of_8tbps")}))


def stats(args):
    """The kernel-trace summary as text: per kernel (short name) its calls,
    total and average duration and share of the process's kernel time."""
    with open(os.path.join(args.dir, "kt", "kt_kernel_stats.csv")) as f:
        rows = list(csv.DictReader(f))
    rows.sort(key=lambda r: -int(r["TotalDurationNs"]))
    total = sum(int(r["TotalDurationNs"]) for r in rows) or 1
    lines = [f"# {args.title}",
             f"# command: rocprofv3 --kernel-trace --stats --output-format csv -d <dir>/kt -o kt "
             f"-- {args.program}; python3 -m scaph_amd.traffic stats --dir <dir> --out {args.out}",
             "# per-kernel totals of the whole process (graph build included)",
             f"{'kernel':<28} {'calls':>7} {'total_ms':>10} {'avg_us':>10} {'pct':>6}"]
    def short(name):
        if "rocprim" in name:
            return "rocprim " + ("radix sort" if "radix_sort" in name else
                                 "partition" if "partition" in name else "primitive")
        return _kernel(name)[:28]
    for r in rows[:args.top]:
        t = int(r["TotalDurationNs"])
        lines.append(f"{short(r['Name']):<28} {int(r['Calls']):>7} {t / 1e6:>10.3f} "
                     f"{float(r['AverageNs']) / 1e3:>10.2f} {100.0 * t / total:>6.2f}")
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("run")
    r.add_argument("--records", required=True)
    r.add_argument("--workload", choices=["rmat", "zipf", "msbfs"], default="rmat")
    r.add_argument("--scale", type=int, default=26)
    r.add_argument("--ef", type=int, default=32)
    r.add_argument("--seed", type=int, default=2)
    r.add_argument("--sources", type=int, default=8)
    r.add_argument("--flag", action="append", default=[])
    r.add_argument("--ms-tune", default="",
                   help="msbfs: this tune (JSON) instead of the autotuner's choice")
    s = sub.add_parser("summarize")
    s.add_argument("--dir", required=True)
    s.add_argument("--out", required=True)
    s.add_argument("--command", default="")
    s.add_argument("--kernel", choices=["k_pull", "k_expand", "k_ms_pull"], default="k_pull")
    t = sub.add_parser("stats")
    t.add_argument("--dir", required=True)
    t.add_argument("--out", required=True)
    t.add_argument("--title", required=True)
    t.add_argument("--program", required=True, help="the profiled command line, for the header")
    t.add_argument("--top", type=int, default=24)
    a = ap.parse_args()
    {"run": run, "summarize": summarize, "stats": stats}[a.cmd](a)


if __name__ == "__main__":
    main()
