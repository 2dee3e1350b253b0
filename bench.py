"""BFS benchmark on one MI355X: R-MAT scale 26, edge factor 32, seed 2.

One "step" is one full BFS from one source (levels final on the device),
sources cycling through 64 seeded sources taken from the giant component.
The graph (paged CSR, out- and in-edge pages) is built on the device
before timing and stays resident in HBM.

Prints ONE JSON line:
  value        median GTEPS over the timed BFS runs (m_c / t_BFS)
  roofline     the dominant kernel (by device time): algorithmic bytes per
               launch / its average launch time (HIP events around it)
  bfs_hbm_frac B_alg / t_BFS / 8 TB/s with B_alg = 4 m_c + 12 n_c + 4 n
  cpu_baseline the oracle's single-core BFS on this host: R-MAT s20 ef16
               seed 1 in full and the first --cpu-sources sources of this graph
"""
import argparse
import ctypes
import hashlib
import json
import os
import statistics
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # bytes/s, MI355X HBM3E


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def pick_sources(eng, n, count, seed, min_reach_frac):
    """count sources, uniformly drawn (seeded) among vertices with out-degree
    >= 1, kept when their BFS reaches >= min_reach_frac of the vertices (the
    giant component); the reach test is an untimed BFS."""
    deg = eng.out_degrees()
    cand = np.nonzero(deg > 0)[0]
    rng = np.random.default_rng(seed)
    order = rng.permutation(len(cand))
    out = []
    tried = 0
    for i in order:
        v = int(cand[i])
        tried += 1
        st = eng.run(v)
        if st["n_reached"] >= min_reach_frac * n:
            out.append(v)
            if len(out) == count:
                break
        if tried > 20 * count:
            break
    return out, tried


def alg_bytes_step(rec, n):
    """Algorithmic bytes of one step's expand kernels:
    pull: Node record 12 B per examined vertex, 4 B per scanned in-edge,
          4 B level write per vertex found, and the visited-read,
          visited-write and next-frontier-write bitmaps (3 * n/8 B);
    push: 12 B Node + 4 B queue read per frontier vertex, 4 B per scanned
          out-edge, 4 B level write + 4 B queue write per vertex reached."""
    if rec["dir"] == 1:
        return 12 * rec["examined"] + 4 * rec["scanned"] + 4 * rec["n_next"] + 3 * (n / 8)
    return 16 * rec["n_f"] + 4 * rec["scanned"] + 8 * rec["n_next"]


def kernel_source_hash():
    """SHA-256 over the kernel library's sources (scaph_amd/csrc and its
    C-ABI header): a traffic profile is used only by the build it was
    measured on."""
    h = hashlib.sha256()
    csrc = os.path.join(ROOT, "scaph_amd", "csrc")
    files = sorted(os.path.join(csrc, f) for f in os.listdir(csrc)
                   if f.endswith((".hip", ".h")))
    for path in files + [os.path.join(ROOT, "include", "scaph_hip.h")]:
        h.update(os.path.basename(path).encode())
        with open(path, "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def device_name(props):
    """The GPU's name as the runtime or rocminfo reports it; where both
    leave it empty (as on some drivers), its architecture, compute units and
    memory."""
    if props.get("name"):
        return props["name"]
    try:
        out = subprocess.run(["rocminfo"], capture_output=True, text=True, timeout=60).stdout
        agent = None
        for line in out.splitlines():
            line = line.strip()
            if line.startswith("Name:"):
                agent = line.split(":", 1)[1].strip()
            if line.startswith("Marketing Name:") and agent and agent.startswith("gfx"):
                name = line.split(":", 1)[1].strip()
                if name:
                    return name
    except (OSError, subprocess.SubprocessError):
        pass
    return (f"{props.get('arch', '')}, {props.get('cu', 0)} CUs, "
            f"{props.get('total_mem', 0) / 2**30:.0f} GiB")


def host_cpu():
    """lscpu's model name and the core counts of this host."""
    model = ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(),
            "cores_allowed": len(os.sched_getaffinity(0))}


def _timed_oracle_bfs(off, ngh, src, runs):
    """Oracle BFS (the reference's check(), app/bfs.cu:67-117) on one pinned
    core: median seconds of `runs` runs and the levels."""
    from tests import _oracle
    cpu = os.sched_getaffinity(0)
    os.sched_setaffinity(0, {min(cpu)})
    try:
        ts = []
        for _ in range(runs):
            t0 = time.perf_counter()
            lv = _oracle.bfs(off, ngh, src)
            ts.append(time.perf_counter() - t0)
    finally:
        os.sched_setaffinity(0, cpu)
    return statistics.median(ts), lv


def cpu_baseline(eng, srcs, runs):
    """The oracle's single-core BFS timed on this host: R-MAT s20 ef16
    seed 1 from its highest out-degree source in full (median of 5), and
    the first len(srcs) sources of this graph, each the median of `runs`
    runs.  `value` is the GTEPS of this graph's sample; the levels of every
    timed source are compared with the GPU's."""
    from tests import _oracle
    c1_off, c1_ngh = _oracle.rmat_csr(20, 16, seed=1)
    c1_deg = np.diff(c1_off.astype(np.int64))
    c1_src = int(np.argmax(c1_deg))
    c1_t, c1_lv = _timed_oracle_bfs(c1_off, c1_ngh, c1_src, 5)
    c1_m = int(c1_deg[c1_lv != 0xFFFFFFFF].sum())
    del c1_off, c1_ngh
    off, ngh = eng.host_csr()
    deg = np.diff(off.astype(np.int64))
    per, parity = [], True
    for s in srcs:
        t, lv = _timed_oracle_bfs(off, ngh, s, runs)
        eng.run(s)
        parity &= bool(np.array_equal(lv, eng.levels()))
        per.append({"src": s, "seconds": t, "m_c": int(deg[lv != 0xFFFFFFFF].sum())})
    m_tot = sum(x["m_c"] for x in per)
    t_tot = sum(x["seconds"] for x in per)
    return {"value": m_tot / t_tot / 1e9, "unit": "GTEPS", "cores": 1, "kind": "port",
            "sample": f"oracle BFS on 1 pinned core: the first {len(srcs)} of the 64 sources of "
                      f"this graph, median of {runs} run(s) each ({t_tot:.1f} s); R-MAT s20 "
                      f"ef16 seed 1 (source {c1_src}) in full, median of 5; levels "
                      f"bit-exact vs GPU: {parity}",
            "per_source": per,
            "rmat_s20_ef16_seed1": {"src": c1_src, "seconds": c1_t, "m_c": c1_m,
                        "gteps": c1_m / c1_t / 1e9, "runs": 5},
            "host": host_cpu()}, parity


def bench_msbfs(scale, ef, queries, seed):
    """Multi-source bitset BFS on R-MAT scale/ef: `queries` sources from the
    giant component in one traversal, autotuned, against the same queries
    run one at a time by the single-source path."""
    import scaph_amd
    eng = scaph_amd.BFS("-max_memory=262144", "-pagesize=10")
    eng.gen_rmat_device(scale, ef, 0.57, 0.19, 0.19, seed=seed)
    info = eng.info()
    srcs, _ = pick_sources(eng, info["n"], queries, seed=seed, min_reach_frac=0.01)
    src = np.array(srcs, np.uint32)
    best, best_ms, log_ = eng.msbfs_autotune(src)
    times = [eng.msbfs(src)["ms"] for _ in range(5)]
    st = eng.msbfs(src)
    single = []
    for s_ in srcs:
        single.append(eng.run(int(s_))["ms"])
    # spot parity: a few queries against the single-source levels
    exact = True
    for q in range(0, len(srcs), max(1, len(srcs) // 8)):
        eng.run(int(srcs[q]))
        exact &= bool(np.array_equal(eng.msbfs_levels(q), eng.levels()))
    ms = statistics.median(times)
    out = {"workload": f"multi-source BFS, R-MAT scale {scale} edge factor {ef} seed {seed}, "
                       f"{len(srcs)} sources from the giant component",
           "queries": len(srcs), "ms": ms, "queries_per_s": len(srcs) / (ms / 1e3),
           "depth": st["depth"], "pairs_reached": st["pairs"],
           "single_source_total_ms": sum(single),
           "speedup_vs_single_source": sum(single) / ms,
           "tune": best, "autotune_candidates": len(log_),
           "persistent_best_ms": min((ms_ for t_, ms_ in log_ if t_.get("persistent")),
                                     default=None),
           "levels_equal_single_source": exact,
           "k_ms_pull_roofline": profiled_roofline("msbfs_s24_ef16_256q_k_ms_pull_traffic.json")}
    eng.close()
    return out


def profiled_roofline(name):
    """Roofline of a kernel from its committed rocprofv3 profile (algorithmic
    bytes per launch over the profiled average launch time, and the
    calibrated FETCH_SIZE + WRITE_SIZE bytes), when the profile was taken on
    this build's kernel sources; None otherwise."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    if t.get("kernel_source_hash") != kernel_source_hash():
        return None
    us = t["avg_launch_us_profiled"]
    alg = t["alg_bytes_per_launch"]
    return {"kernel": t["kernel"], "bound": "hbm", "avg_launch_us": us,
            "alg_bytes_per_launch": alg, "achieved": alg / (us * 1e-6) / 1e9, "peak": 8000.0,
            "unit": "GB/s", "frac": alg / (us * 1e-6) / 8e12,
            "traffic": t["traffic_bytes_per_launch"], "launches": t["launches"],
            "profile": os.path.join("profiles", name)}


def bench_small_batch(count, seed):
    """4096 R-MAT graphs (scale 10-16, edge factor 8-32), one query each, in
    one persistent launch; a sample checked against the oracle."""
    import scaph_amd
    b = scaph_amd.Batch(count=count, min_scale=10, max_scale=16, min_ef=8, max_ef=32, seed=seed)
    info = b.info()
    best, best_ms, log_ = b.autotune()
    times = [b.run() for _ in range(5)]
    ms = statistics.median(times)
    exact = None
    try:
        sys.path.insert(0, ROOT)
        from tests import _oracle as O
        exact = True
        for g in range(0, count, max(1, count // 16)):
            off, ngh = b.csr(g)
            exact &= bool(np.array_equal(b.levels(g), O.bfs(off, ngh, b.graph(g)["src"])))
    except ImportError:
        pass
    out = {"workload": f"{count} R-MAT graphs, scale 10-16, edge factor 8-32, seed {seed}, "
                       "one source each, one persistent launch",
           "graphs": info["graphs"], "vertices": info["vertices"], "edges": info["edges"],
           "ms": ms, "queries_per_s": count / (ms / 1e3),
           "gteps": info["edges"] / (ms / 1e3) / 1e9,
           "tune": best, "autotune_candidates": len(log_),
           "levels_exact_vs_oracle_sample": exact}
    b.close()
    return out


SSSP_INF = 0x7FFFFFFF  # unreached distance (MAXPATH)


def bench_apps(scale, ef, seed, wmax=255):
    """The WCC and SSSP apps on the resident R-MAT scale/ef graph (weights
    uniform on [1, wmax] for SSSP), timed; full-size results checked by the
    properties that fix them: the SSSP distances satisfy d[v] <= d[u] + w on
    every edge and every reached v != src has a tight in-edge (together they
    characterise shortest paths for positive weights); the WCC labels are a
    fixpoint (label[v] <= label[u] on every edge u->v, label[v] <= v)."""
    import scaph_amd
    eng = scaph_amd.BFS("-max_memory=262144", "-pagesize=10")
    eng.gen_rmat_device(scale, ef, 0.57, 0.19, 0.19, seed=seed, keep_host=True)
    off, ngh = eng.host_csr()
    n, m = len(off) - 1, len(ngh)
    rng = np.random.default_rng(seed)
    wgh = rng.integers(1, wmax + 1, m, dtype=np.uint32)
    eng.set_weights(wgh)
    src_of = np.repeat(np.arange(n, dtype=np.uint32), np.diff(off.astype(np.int64)))
    out = {"workload": f"R-MAT scale {scale} edge factor {ef} seed {seed}, resident; "
                       f"SSSP weights uniform on [1, {wmax}]", "vertices": n, "edges": m}
    wcc = [eng.wcc() for _ in range(3)]
    lab = eng.levels()
    ms = statistics.median(r["ms"] for r in wcc)
    out["wcc"] = {"ms": ms, "gteps": m / (ms / 1e3) / 1e9, "rounds": wcc[-1]["depth"],
                  "edges_relaxed": wcc[-1]["m_reached"],
                  "components": int(np.count_nonzero(lab == np.arange(n, dtype=np.uint32))),
                  "fixpoint": bool((lab[ngh] <= lab[src_of]).all()
                                   and (lab <= np.arange(n, dtype=np.uint32)).all())}
    del lab
    srcs, _ = pick_sources(eng, n, 3, seed=seed, min_reach_frac=0.01)
    runs, exact = [], True
    for s_ in srcs:
        runs.append(eng.sssp(s_))
        d = eng.levels().astype(np.int64)
        du = d[src_of]
        reach_u = du < SSSP_INF
        cand = du + wgh
        ok = bool((d[ngh][reach_u] <= cand[reach_u]).all())
        tight = np.zeros(n, bool)
        tight[ngh[reach_u & (cand == d[ngh])]] = True
        reached = d < SSSP_INF
        tight[s_] = True
        ok &= bool(d[s_] == 0 and tight[reached].all())
        exact &= ok
    ms = statistics.median(r["ms"] for r in runs)
    out["sssp"] = {"ms": ms, "gteps": m / (ms / 1e3) / 1e9, "sources": srcs,
                   "rounds": [r["depth"] for r in runs],
                   "edges_relaxed": [r["m_reached"] for r in runs],
                   "shortest_path_conditions_hold": exact}
    eng.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--scale", type=int, default=26)
    ap.add_argument("--ef", type=int, default=32)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--sources", type=int, default=64)
    ap.add_argument("--schedule", default="fixed")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--flag", action="append", default=[],
                    help="extra engine flag, e.g. --flag=-pull_summary=false")
    ap.add_argument("--no-batched", dest="batched", action="store_false",
                    help="skip the batched BFS measurements")
    ap.add_argument("--no-apps", dest="apps", action="store_false",
                    help="skip the WCC and SSSP measurements")
    ap.add_argument("--no-learned", dest="learned", action="store_false",
                    help="skip the learned-schedule pass")
    ap.add_argument("--learned-check", type=int, default=8,
                    help="sources whose levels are compared between the schedules")
    ap.add_argument("--cpu-sources", type=int, default=1,
                    help="sources of this graph the CPU baseline times (the full protocol: 4)")
    ap.add_argument("--cpu-runs", type=int, default=1,
                    help="runs per CPU-baseline source, median taken (the full protocol: 5)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="run one device-to-device copy of this size first (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the BFS here has no multi-GPU path; the per-GPU "
              "work at full shape is what --gpus 1 runs", file=sys.stderr)
        sys.exit(2)

    import scaph_amd
    from scaph_amd import _lib

    props = scaph_amd.device_props(0)
    log("device", props)
    eng = scaph_amd.BFS("-max_memory=262144", "-pagesize=10", f"-schedule={args.schedule}",
                        *args.flag)
    t0 = time.perf_counter()
    eng.gen_rmat_device(args.scale, args.ef, 0.57, 0.19, 0.19, seed=args.seed,
                        keep_host=not args.no_cpu)
    info = eng.info()
    log("graph built", info, f"{time.perf_counter() - t0:.1f}s")
    n = info["n"]

    if args.calib_copy_mib:
        nbytes = args.calib_copy_mib << 20
        hip = _lib.hip()
        a, b = ctypes.c_void_p(), ctypes.c_void_p()
        _lib.check(hip.scaph_dev_alloc(ctypes.byref(a), nbytes))
        _lib.check(hip.scaph_dev_alloc(ctypes.byref(b), nbytes))
        _lib.check(hip.scaph_copy_d2d_async(b, a, nbytes, None))
        _lib.check(hip.scaph_device_sync())
        hip.scaph_dev_free(a)
        hip.scaph_dev_free(b)
        log(f"calibration copy: {nbytes} bytes device-to-device")

    srcs, tried = pick_sources(eng, n, args.sources, seed=args.seed, min_reach_frac=0.01)
    log(f"sources: {len(srcs)} of {tried} candidates reach >=1% of V")
    if not srcs:
        raise SystemExit("no source reaches 1% of the vertices")

    for i in range(args.warmup):
        eng.run(srcs[i % len(srcs)])

    hip = _lib.hip()
    _lib.check(hip.scaph_device_sync())
    runs = []
    t_start = time.perf_counter()
    for k in range(args.steps):
        st = eng.run(srcs[k % len(srcs)])
        st["records"] = eng.records()
        runs.append(st)
    _lib.check(hip.scaph_device_sync())
    wall = time.perf_counter() - t_start

    gteps = [r["m_reached"] / (r["ms"] / 1e3) / 1e9 for r in runs]
    mid = sorted(range(len(runs)), key=lambda i: gteps[i])[len(runs) // 2]
    for rec in runs[mid]["records"]:
        log("level", rec["level"], "dir", "pull" if rec["dir"] else "push", "n_f", rec["n_f"],
            "m_f", rec["m_f"], "n_next", rec["n_next"], "scanned", rec["scanned"],
            "examined", rec["examined"], f"kernel_ms {rec['kernel_ms']:.3f}",
            f"step_ms {rec['ms']:.3f}")
    log(f"median run: t_bfs {runs[mid]['ms']:.3f} ms, n_c {runs[mid]['n_reached']}, "
        f"m_c {runs[mid]['m_reached']}")
    balg = [4 * r["m_reached"] + 12 * r["n_reached"] + 4 * n for r in runs]
    frac = [b / (r["ms"] / 1e3) / HBM_PEAK for b, r in zip(balg, runs)]
    ref_gteps = [info["m"] / (r["ms"] / 1e3) / 1e9 for r in runs]

    # dominant kernel by summed kernel time: the push expand (0), the pull
    # over windows (1, k_pull) or the pull over a list of the few unvisited
    # vertices left (2, k_pull_sparse)
    agg = {0: [0.0, 0, 0.0], 1: [0.0, 0, 0.0], 2: [0.0, 0, 0.0]}  # -> [ms, launches, alg bytes]
    for r in runs:
        for rec in r["records"]:
            a = agg[2 if rec["dir"] == 1 and rec.get("sparse") else rec["dir"]]
            a[0] += rec["kernel_ms"]
            a[1] += 1
            a[2] += alg_bytes_step(rec, n)
    dom = max(agg, key=lambda d: agg[d][0])
    ms_tot, launches, bytes_tot = agg[dom]
    avg_ms = ms_tot / max(launches, 1)
    avg_bytes = bytes_tot / max(launches, 1)
    achieved = avg_bytes / (avg_ms / 1e3) / 1e9 if avg_ms > 0 else 0.0
    traffic = None
    traffic_src = None
    tfile = os.path.join(ROOT, "profiles", "bfs_s26_ef32_k_pull_traffic.json")
    if dom == 1 and os.path.exists(tfile) and args.scale == 26 and args.ef == 32:
        # HBM-side bytes per k_pull launch from rocprofv3 FETCH_SIZE/WRITE_SIZE
        # passes over this workload, taken only when they were measured on
        # the kernel sources of this build
        with open(tfile) as f:
            tj = json.load(f)
        if tj.get("kernel_source_hash") == kernel_source_hash():
            traffic = tj.get("traffic_bytes_per_launch")
            traffic_src = os.path.relpath(tfile, ROOT)
    # the profiled bytes over this run's launch time: how hard the kernel
    # drives the memory side, beside the algorithmic figure above (counters
    # include Infinity Cache hits, so this bounds HBM bandwidth from above)
    t_gbps = traffic / (avg_ms / 1e3) / 1e9 if traffic and avg_ms > 0 else None
    roofline = {"bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                "frac": achieved / 8000.0, "traffic": traffic, "traffic_profile": traffic_src,
                "traffic_achieved": t_gbps,
                "traffic_frac": t_gbps / 8000.0 if t_gbps else None,
                "kernel": {0: "push (k_prep_queue+k_scan_blocks+k_mark_tiles+k_expand)",
                           1: "k_pull", 2: "k_pull_sparse"}[dom],
                "avg_launch_ms": avg_ms, "alg_bytes_per_launch": avg_bytes, "launches": launches,
                "share_of_kernel_time": ms_tot / max(sum(a[0] for a in agg.values()), 1e-9),
                "pull_sparse": {"avg_launch_ms": agg[2][0] / max(agg[2][1], 1),
                                "launches": agg[2][1],
                                "alg_bytes_per_launch": agg[2][2] / max(agg[2][1], 1)}}

    # the learned schedule (committed model) on the same sources, paired
    # with the fixed schedule source by source (fixed, then learned, on the
    # same engine, one after the other) so that clock and cache drift between
    # phases cannot enter the ratio; levels compared with the fixed
    # schedule's for the first learned_check sources
    learned = None
    model_path = os.path.join(ROOT, "scaph_amd", "schedule_mlp.bin")
    if args.learned and os.path.exists(model_path):
        from scaph_amd import learn as L
        L.model_for_graph(eng, model_path)
        fixed_levels = {}
        for src in srcs[:args.learned_check]:
            eng.set_schedule("fixed")
            eng.run(src)
            fixed_levels[src] = eng.levels()
        for i in range(args.warmup):
            for sched in ("fixed", "learned"):
                eng.set_schedule(sched)
                eng.run(srcs[i % len(srcs)])
        _lib.check(hip.scaph_device_sync())
        fruns, lruns = [], []
        for k in range(args.steps):
            for sched, out in (("fixed", fruns), ("learned", lruns)):
                eng.set_schedule(sched)
                st = eng.run(srcs[k % len(srcs)])
                st["dirs"] = [r["dir"] for r in eng.records()]
                out.append(st)
        _lib.check(hip.scaph_device_sync())
        equal = True
        eng.set_schedule("learned")
        for src, want in fixed_levels.items():
            eng.run(src)
            equal &= bool(np.array_equal(eng.levels(), want))
        eng.set_schedule("fixed")
        lg = [r["m_reached"] / (r["ms"] / 1e3) / 1e9 for r in lruns]
        lt = statistics.median(r["ms"] for r in lruns)
        ft = statistics.median(r["ms"] for r in fruns)
        same_dirs = sum(1 for a, b in zip(fruns, lruns) if a["dirs"] == b["dirs"])
        learned = {"gteps_median": statistics.median(lg), "t_bfs_ms_median": lt,
                   "fixed_t_bfs_ms_median_paired": ft,
                   "speedup_vs_fixed": ft / lt,
                   "speedup_vs_fixed_total": sum(r["ms"] for r in fruns) /
                                             sum(r["ms"] for r in lruns),
                   "levels_equal_fixed": equal, "sources_checked": len(fixed_levels),
                   "runs_with_same_directions": same_dirs, "paired_runs": len(lruns),
                   "learned_margin": next((float(f.split("=", 1)[1]) for f in args.flag
                                           if f.lstrip("-").startswith("learned_margin=")),
                                          0.5),
                   "model": "scaph_amd/schedule_mlp.bin"}
        log("learned schedule", learned)

    # parity at full size: device invariant check on the last source
    vsrc = runs[-1]["src"]
    eng.set_schedule("fixed")
    eng.run(vsrc)  # the level array must belong to the source being validated
    viol = eng.validate(vsrc)
    cpu = None
    parity = None
    if not args.no_cpu:
        cpu, parity = cpu_baseline(eng, srcs[:args.cpu_sources], args.cpu_runs)
        log("cpu baseline", cpu)

    batched = None
    if args.batched:
        eng.close()
        batched = {}
        log("batched: multi-source")
        batched["multi_source"] = bench_msbfs(24, 16, 256, seed=3)
        log("batched:", batched["multi_source"])
        log("batched: small graphs")
        batched["small_graphs"] = bench_small_batch(4096, seed=4)
        log("batched:", batched["small_graphs"])

    apps = None
    if args.apps:
        if not args.batched:
            eng.close()
        log("apps: WCC and SSSP")
        apps = bench_apps(24, 16, seed=5)
        log("apps:", apps)

    med = statistics.median(gteps)
    out = {
        "metric": "BFS GTEPS, R-MAT s26 ef32, learned vs fixed-alpha schedule; % HBM peak",
        "value": med,
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic R-MAT built on the device (SplitMix64 stream, seeded relabelling)",
        "config": {"workload": f"BFS R-MAT scale {args.scale} edge factor {args.ef} "
                               f"(a,b,c)=(.57,.19,.19) seed {args.seed}; {len(srcs)} "
                               f"giant-component sources, {args.steps} timed BFS runs over "
                               f"sources 0..{min(args.steps, len(srcs)) - 1} in order",
                   "sources": len(srcs), "sources_timed": min(args.steps, len(srcs)),
                   "vertices": n, "edges": info["m"], "pages": info["pages"],
                   "schedule": args.schedule, "alpha": 14, "beta": 24},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "learned": learned,
        "batched": batched,
        "apps": apps,
        "bfs_hbm_frac_median": statistics.median(frac),
        "t_bfs_ms_median": statistics.median(r["ms"] for r in runs),
        "gteps_ref_formula_median": statistics.median(ref_gteps),
        "gteps_mean_of_runs": sum(r["m_reached"] for r in runs) / sum(r["ms"] for r in runs) / 1e6,
        "levels_exact_vs_oracle": parity,
        "invariant_violations": list(viol),
        "device": device_name(props),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
