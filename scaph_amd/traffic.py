"""Memory-side traffic of the BFS kernels, from rocprofv3 counters.

Two steps:

  rocprofv3 --pmc FETCH_SIZE -d <dir>/fetch -o fetch --output-format csv -- \\
      python3 -m scaph_amd.traffic run --records <dir>/records.json
  rocprofv3 --pmc WRITE_SIZE -d <dir>/write -o write --output-format csv -- \\
      python3 -m scaph_amd.traffic run --records <dir>/records.json
  python3 -m scaph_amd.traffic summarize --dir <dir> --out profiles/<name>.json

`run` first reads and writes a buffer of known size with the library's own
kernels (scaph_checksum_u32 reads 4n bytes, scaph_fill_u32 writes 4n bytes),
then builds the benchmark graph (R-MAT scale 26, edge factor 32, seed 2 by
default) and runs BFS from its giant-component sources, writing every run's
per-level records.  `summarize` scales FETCH_SIZE and WRITE_SIZE (KiB per
dispatch) by what they read for those two kernels, maps the k_pull
dispatches of the timed runs onto their pull levels, and writes bytes per
launch beside the algorithmic bytes of the same launches, stamped with the
hash of the kernel sources (bench.py uses the figure only for that build).
"""
import argparse
import csv
import ctypes
import json
import os
import re
import statistics
import sys

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
    flags = ["-max_memory=262144", "-pagesize=10"] + args.flag
    e = scaph_amd.BFS(*flags)
    e.gen_rmat_device(args.scale, args.ef, 0.57, 0.19, 0.19, seed=args.seed)
    info = e.info()
    srcs, _ = bench.pick_sources(e, info["n"], args.sources, seed=args.seed, min_reach_frac=0.01)
    runs = []
    for s in srcs[:2] + srcs:  # two warm-up runs, then the timed ones
        st = e.run(s)
        runs.append({"src": s, "ms": st["ms"], "n_reached": st["n_reached"],
                     "m_reached": st["m_reached"], "levels": e.records()})
    e.close()
    with open(args.records, "w") as f:
        json.dump({"n": info["n"], "m": info["m"], "flags": flags, "timed_from": 2,
                   "runs": runs}, f)
    print(json.dumps({"runs": len(runs), "sources": len(srcs)}), flush=True)


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


def pull_alg_bytes(rec, n):
    """Algorithmic bytes of one pull launch (as bench.py counts them)."""
    return 12 * rec["examined"] + 4 * rec["scanned"] + 4 * rec["n_next"] + 3 * (n / 8)


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
    # pull levels of the timed runs, in launch order
    timed = rec["runs"][rec["timed_from"]:]
    pull_levels = [lv for r in timed for lv in r["levels"] if lv["dir"] == 1]
    fk = [x for x in fetch if x["kernel"] == "k_pull"]
    wk = [x for x in write if x["kernel"] == "k_pull"]
    if len(fk) != len(wk) or len(fk) < len(pull_levels):
        raise SystemExit(f"k_pull dispatches: {len(fk)} fetch, {len(wk)} write, "
                         f"{len(pull_levels)} timed pull levels")
    fk, wk = fk[-len(pull_levels):], wk[-len(pull_levels):]
    per = []
    for lv, a, b in zip(pull_levels, fk, wk):
        fb = a["kib"] * 1024 * f_scale
        wb = b["kib"] * 1024 * w_scale
        per.append({"level": lv["level"], "examined": lv["examined"], "scanned": lv["scanned"],
                    "n_next": lv["n_next"], "fetch_bytes": fb, "write_bytes": wb,
                    "alg_bytes": pull_alg_bytes(lv, rec["n"]), "us_profiled": a["us"]})
    k = len(per)
    traffic = sum(p["fetch_bytes"] + p["write_bytes"] for p in per) / k
    alg = sum(p["alg_bytes"] for p in per) / k
    out = {"command": args.command, "kernel": "k_pull", "launches": k,
           "kernel_source_hash": bench.kernel_source_hash(),
           "calibration": {"read_kernel": "k_checksum_u32", "write_kernel": "k_fill_u32",
                           "bytes": calib_bytes, "fetch_raw_bytes": rd[0]["kib"] * 1024,
                           "write_raw_bytes": max(wr, key=lambda x: x["kib"])["kib"] * 1024,
                           "fetch_scale": f_scale, "write_scale": w_scale},
           "fetch_bytes_per_launch": sum(p["fetch_bytes"] for p in per) / k,
           "write_bytes_per_launch": sum(p["write_bytes"] for p in per) / k,
           "traffic_bytes_per_launch": traffic, "alg_bytes_per_launch": alg,
           "traffic_over_alg": traffic / alg if alg else None,
           "avg_launch_us_profiled": statistics.mean(p["us_profiled"] for p in per),
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
    print(json.dumps({k_: out[k_] for k_ in ("launches", "fetch_bytes_per_launch",
                                             "write_bytes_per_launch", "alg_bytes_per_launch",
                                             "traffic_over_alg")}))


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("run")
    r.add_argument("--records", required=True)
    r.add_argument("--scale", type=int, default=26)
    r.add_argument("--ef", type=int, default=32)
    r.add_argument("--seed", type=int, default=2)
    r.add_argument("--sources", type=int, default=8)
    r.add_argument("--flag", action="append", default=[])
    s = sub.add_parser("summarize")
    s.add_argument("--dir", required=True)
    s.add_argument("--out", required=True)
    s.add_argument("--command", default="")
    a = ap.parse_args()
    run(a) if a.cmd == "run" else summarize(a)


if __name__ == "__main__":
    main()
