"""Learned direction schedule: trace collection, training on the GPU with the
project's own MLP kernels (scaph_mlp_forward / _backward / _sgd), the model
file, and the A/B evaluation against the fixed alpha/beta rule.

Training data are per-level traces of plain BFS runs over seeded graphs
(R-MAT scales 18-24 with a in [0.45, 0.65], Erdos-Renyi with average degree
8-64; --powerlaw adds power-law digraphs with Pareto out-degrees and uniform
destinations): at each level both directions run from the same state, the step times
become a soft label (is pull faster?) and the fraction of pages active at the
next level is the second target.  Levels never depend on the model; only the
wall time does.

    python -m scaph_amd.learn --graphs 200 --out scaph_amd/schedule_mlp.bin
"""
import argparse
import ctypes
import json
import os
import struct
import sys
import time

import numpy as np

from . import BFS, _lib

NPARAM = 64 * 16 + 64 + 64 * 64 + 64 + 2 * 64 + 2
HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL = os.path.join(HERE, "schedule_mlp.bin")


def log(*a):
    print("[learn]", *a, file=sys.stderr, flush=True)


# ----------------------------------------------------------------- graphs

def training_specs(count=200, seed=5, powerlaw=False):
    """Seeded training graphs: 70% R-MAT (scale 18-24, a in [0.45, 0.65],
    b = c = 0.4(1-a), edge factor 8-32), 30% Erdos-Renyi (2^18-2^21
    vertices, average degree 8-64) -- the committed model's set.

    powerlaw: 60% R-MAT, 20% Erdos-Renyi, 20% power-law digraphs (2^18-2^21
    vertices, Pareto out-degrees with x_min in [1, 2] and alpha in
    [1.7, 2.1] capped at n/100, uniform destinations)."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        k = i % 10
        if k < (6 if powerlaw else 7):
            a = float(rng.uniform(0.45, 0.65))
            specs.append({"kind": "rmat", "scale": int(rng.integers(18, 25)),
                          "ef": int(rng.integers(8, 33)), "a": a, "b": 0.4 * (1 - a),
                          "c": 0.4 * (1 - a), "seed": 1000 + i})
        elif not powerlaw or k < 8:
            n = 1 << int(rng.integers(18, 22))
            specs.append({"kind": "er", "n": n, "m": n * int(rng.integers(8, 65)),
                          "seed": 2000 + i})
        else:
            n = 1 << int(rng.integers(18, 22))
            specs.append({"kind": "zipf", "n": n, "x_min": float(rng.uniform(1.0, 2.0)),
                          "alpha": float(rng.uniform(1.7, 2.1)), "cap": max(1000, n // 100),
                          "seed": 3000 + i})
    return specs


def build(eng, spec):
    if spec["kind"] == "rmat":
        eng.gen_rmat_device(spec["scale"], spec["ef"], spec["a"], spec["b"], spec["c"],
                            seed=spec["seed"])
    elif spec["kind"] == "er":
        eng.gen_er_device(spec["n"], spec["m"], seed=spec["seed"])
    else:
        eng.gen_zipf(spec["n"], spec["x_min"], spec["alpha"], spec["cap"], seed=spec["seed"])


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


def collect(specs, sources=2, reps=2):
    eng = BFS("-max_memory=262144")
    recs = []
    t0 = time.time()
    for gi, spec in enumerate(specs):
        build(eng, spec)
        inf = eng.info()
        for src in pick_sources(eng, sources, seed=spec["seed"]):
            for r in eng.explore(src, reps):
                r["graph"] = gi
                r["n"], r["m"] = inf["n"], inf["m"]
                recs.append(r)
        if gi % 10 == 0:
            log(f"traces: graph {gi}/{len(specs)} ({spec['kind']}), {len(recs)} records, "
                f"{time.time() - t0:.0f}s")
    eng.close()
    return recs


def features_labels(recs):
    X = np.array([r["x"] for r in recs], np.float32)
    tp = np.array([r["t_push_ms"] for r in recs], np.float64)
    tq = np.array([r["t_pull_ms"] for r in recs], np.float64)
    tau = 0.1 * np.maximum(np.maximum(tp, tq), 1e-3)
    t0 = 1.0 / (1.0 + np.exp(-(tp - tq) / tau))  # soft "pull is faster"
    t1 = np.array([r["pages_active_next"] for r in recs], np.float32)
    return X, np.stack([t0.astype(np.float32), t1], 1)


# ------------------------------------------------------------ device MLP

class DevArray:
    def __init__(self, arr=None, nbytes=None):
        hip = _lib.hip()
        self.nbytes = int(arr.nbytes if arr is not None else nbytes)
        self.ptr = ctypes.c_void_p()
        _lib.check(hip.scaph_dev_alloc(ctypes.byref(self.ptr), max(self.nbytes, 8)))
        if arr is not None:
            a = np.ascontiguousarray(arr)
            _lib.check(hip.scaph_copy_h2d_async(self.ptr, a.ctypes.data, a.nbytes, None))
            _lib.check(hip.scaph_device_sync())

    def get(self, dtype, count):
        out = np.empty(count, dtype)
        hip = _lib.hip()
        _lib.check(hip.scaph_copy_d2h_async(out.ctypes.data, self.ptr, out.nbytes, None))
        _lib.check(hip.scaph_device_sync())
        return out

    def put(self, arr):
        a = np.ascontiguousarray(arr)
        hip = _lib.hip()
        _lib.check(hip.scaph_copy_h2d_async(self.ptr, a.ctypes.data, a.nbytes, None))
        _lib.check(hip.scaph_device_sync())

    def __del__(self):
        try:
            _lib.hip().scaph_dev_free(self.ptr)
        except Exception:
            pass


def f32_to_bf16(x):
    """Round-to-nearest-even onto the bf16 grid (bit patterns as uint16)."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    nan = (u & 0x7F800000) == 0x7F800000
    nan &= (u & 0x007FFFFF) != 0
    r[nan] = ((u[nan] >> 16) | 0x40).astype(np.uint16)
    return r


def bf16_to_f32(h):
    return (np.asarray(h, np.uint16).astype(np.uint32) << 16).view(np.float32)


SCALAR_FEATURES = 4  # log2(1+n_f), log2(1+m_f), log2(1+m_u), level
RELATIVE = 1  # model flag: features 0..2 are taken relative to log2 n, log2 m, log2 m


def relative_features(X, n, m):
    """Features 0..2 as log2 of the frontier's share of the graph: the
    device kernels compute log2(1+n_f), log2(1+m_f), log2(1+m_u); a model
    trained on these minus log2 n, log2 m, log2 m compares shares, as the
    alpha/beta rule does, and so carries across graph sizes."""
    X = np.array(X, np.float32, copy=True)
    X[:, 0] -= np.log2(np.asarray(n, np.float64)).astype(np.float32)
    X[:, 1] -= np.log2(np.asarray(m, np.float64)).astype(np.float32)
    X[:, 2] -= np.log2(np.asarray(m, np.float64)).astype(np.float32)
    return X


def graph_mean(mean, flags, n, m):
    """The feature offsets the device uses for a graph of n vertices and m
    edges: (x - offset) / scale with offset = mean, plus log2 n / log2 m on
    features 0..2 for a relative model (normalisation is affine, so the
    kernels need no change)."""
    mean = np.array(mean, np.float32, copy=True)
    if flags & RELATIVE:
        mean[0] += np.float32(np.log2(max(n, 1)))
        mean[1] += np.float32(np.log2(max(m, 1)))
        mean[2] += np.float32(np.log2(max(m, 1)))
    return mean


def init_params(seed=0, scalar=False):
    """He-normal weights, zero biases.  scalar: W1's columns for the 12
    histogram features start (and, with those inputs held at zero, stay)
    zero, so the model reads the scalar features only."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    o = 0
    for fan_in, fan_out in ((16, 64), (64, 64), (64, 2)):
        w = rng.normal(0, np.sqrt(2.0 / fan_in), (fan_out, fan_in)).astype(np.float32)
        if scalar and fan_in == 16:
            w[:, SCALAR_FEATURES:] = 0.0
        p[o:o + w.size] = w.ravel()
        o += w.size + fan_out  # biases stay 0
    return p


class DeviceMLP:
    """Parameters, optimiser state and one batch on the device."""

    def __init__(self, master):
        self.master = DevArray(np.asarray(master, np.float32))
        self.mom = DevArray(np.zeros(NPARAM, np.float32))
        self.params = DevArray(f32_to_bf16(master))
        self.grad = DevArray(nbytes=NPARAM * 4)
        self.loss = DevArray(nbytes=4)

    def forward(self, X, keep_act=False):
        hip = _lib.hip()
        B = len(X)
        dx = DevArray(np.asarray(X, np.float32))
        dy = DevArray(nbytes=B * 8)
        act = DevArray(nbytes=B * 128 * 4) if keep_act else None
        _lib.check(hip.scaph_mlp_forward(self.params.ptr, dx.ptr, B, dy.ptr,
                                         act.ptr if act else None, None))
        return dx, dy, act

    def step(self, dx, dT, B, lr, mu, w1):
        hip = _lib.hip()
        if getattr(self, "_bufB", None) != B:
            self._dy = DevArray(nbytes=B * 8)
            self._act = DevArray(nbytes=B * 128 * 4)
            self._bufB = B
        dy, act = self._dy, self._act
        _lib.check(hip.scaph_mlp_forward(self.params.ptr, dx.ptr, B, dy.ptr, act.ptr, None))
        _lib.check(hip.scaph_mlp_backward(self.params.ptr, dx.ptr, act.ptr, dy.ptr, dT.ptr, B, w1,
                                          self.grad.ptr, self.loss.ptr, None))
        _lib.check(hip.scaph_mlp_sgd(self.master.ptr, self.mom.ptr, self.grad.ptr, NPARAM, lr, mu,
                                     self.params.ptr, None))
        return self.loss

    def params_bf16(self):
        return self.params.get(np.uint16, NPARAM)

    def predict(self, X):
        _, dy, _ = self.forward(X)
        return dy.get(np.float32, 2 * len(X)).reshape(-1, 2)


def train(X, T, steps=4000, lr=0.05, mu=0.9, w1=1.0, seed=0, scalar=False):
    if scalar:
        X = np.array(X, np.float32)
        X[:, SCALAR_FEATURES:] = 0.0
    mean = X.mean(0)
    scale = X.std(0)
    scale[scale < 1e-6] = 1.0
    Xn = ((X - mean) / scale).astype(np.float32)
    mlp = DeviceMLP(init_params(seed, scalar))
    dx = DevArray(Xn)
    dT = DevArray(np.asarray(T, np.float32))
    losses = []
    for s in range(steps):
        loss = mlp.step(dx, dT, len(Xn), lr, mu, w1)
        if s % 500 == 0 or s == steps - 1:
            losses.append((s, float(loss.get(np.float32, 1)[0])))
    return mlp, mean.astype(np.float32), scale.astype(np.float32), losses


# -------------------------------------------------------------- model file

def save_model(path, params_u16, mean, scale, gain=0.0, flags=0):
    """Raw little-endian: NPARAM u16 parameters, 16 f32 means, 16 f32
    scales, one f32 prefetch gain, one u32 of flags (RELATIVE; files
    without it have absolute features) -- the layout
    LearnedSchedule::load_file reads."""
    with open(path, "wb") as f:
        f.write(np.asarray(params_u16, np.uint16).tobytes())
        f.write(np.asarray(mean, np.float32).tobytes())
        f.write(np.asarray(scale, np.float32).tobytes())
        f.write(struct.pack("<f", float(gain)))
        f.write(struct.pack("<I", int(flags)))


def load_model(path=DEFAULT_MODEL):
    raw = open(path, "rb").read()
    p = np.frombuffer(raw[:NPARAM * 2], np.uint16)
    o = NPARAM * 2
    mean = np.frombuffer(raw[o:o + 64], np.float32)
    scale = np.frombuffer(raw[o + 64:o + 128], np.float32)
    gain = struct.unpack("<f", raw[o + 128:o + 132])[0]
    return p.copy(), mean.copy(), scale.copy(), gain


def load_model_flags(path=DEFAULT_MODEL):
    raw = open(path, "rb").read()
    o = NPARAM * 2 + 132
    return struct.unpack("<I", raw[o:o + 4])[0] if len(raw) >= o + 4 else 0


def model_for_graph(eng, path=DEFAULT_MODEL):
    """Loads the model into engine `eng` with the offsets of its loaded graph."""
    p, mean, scale, _ = load_model(path)
    inf = eng.info()
    eng.set_model(p, graph_mean(mean, load_model_flags(path), inf["n"], inf["m"]), scale)


# ------------------------------------------------------------- evaluation

def compare(eng, sources, repeats=3):
    """Median t_BFS per source for the fixed rule and the learned model on
    the loaded graph; levels must agree bit for bit."""
    out = []
    for src in sources:
        ms = {}
        lv = {}
        for sched in ("fixed", "learned"):
            eng.set_schedule(sched)
            t = []
            for _ in range(repeats):
                t.append(eng.run(src)["ms"])
            ms[sched] = float(np.median(t))
            lv[sched] = eng.levels()
        if not np.array_equal(lv["fixed"], lv["learned"]):
            raise AssertionError(f"levels differ between schedules for source {src}")
        out.append({"src": src, "fixed_ms": ms["fixed"], "learned_ms": ms["learned"]})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--graphs", type=int, default=200)
    ap.add_argument("--sources", type=int, default=2)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--out", default=DEFAULT_MODEL)
    ap.add_argument("--traces", default="", help="also save the training traces here (.npz)")
    ap.add_argument("--load-traces", default="", help="train on saved traces instead of new ones")
    ap.add_argument("--absolute", action="store_true",
                    help="train on absolute log sizes (the older models) instead of shares")
    ap.add_argument("--powerlaw", action="store_true",
                    help="train on R-MAT, Erdos-Renyi and power-law graphs (60/20/20)")
    ap.add_argument("--histogram", action="store_true",
                    help="let the model read the 12 page-histogram features too (the decision "
                         "then needs a frontier pass per level)")
    args = ap.parse_args()
    scalar = not args.histogram
    if args.load_traces:
        z = np.load(args.load_traces, allow_pickle=False)
        X, T = z["X"], z["T"]
        if "n" in z:
            gn, gm = z["n"], z["m"]
        else:
            # traces saved without the sizes: n from the (seeded) specs, m
            # from each graph's first record, m_u + m_f of the source
            specs = training_specs(args.graphs, powerlaw=args.powerlaw)
            g = z["graph"]
            gn = np.array([2 ** specs[i]["scale"] if specs[i]["kind"] == "rmat"
                           else specs[i]["n"] for i in g], np.float64)
            first = {}
            for k, i in enumerate(g):
                first.setdefault(int(i), k)
            mm = {i: (2.0 ** X[k, 2] - 1) + (2.0 ** X[k, 1] - 1) for i, k in first.items()}
            gm = np.array([mm[int(i)] for i in g], np.float64)
        recs = None
    else:
        specs = training_specs(args.graphs, powerlaw=args.powerlaw)
        recs = collect(specs, args.sources)
        X, T = features_labels(recs)
        gn = np.array([r["n"] for r in recs], np.float64)
        gm = np.array([r["m"] for r in recs], np.float64)
    if args.traces and recs is not None:
        np.savez_compressed(args.traces, X=X, T=T, n=gn, m=gm,
                            t_push=np.array([r["t_push_ms"] for r in recs], np.float32),
                            t_pull=np.array([r["t_pull_ms"] for r in recs], np.float32),
                            graph=np.array([r["graph"] for r in recs], np.int32))
    flags = 0 if args.absolute else RELATIVE
    if flags & RELATIVE:
        X = relative_features(X, gn, gm)
    log(f"{len(X)} training records; pull faster in {(T[:, 0] > 0.5).mean() * 100:.1f}%")
    mlp, mean, scale, losses = train(X, T, steps=args.steps, scalar=scalar)
    log("loss", losses)
    Xs = np.array(X, np.float32)
    if scalar:
        Xs[:, SCALAR_FEATURES:] = 0.0
    pred = mlp.predict(((Xs - mean) / scale).astype(np.float32))
    acc = ((pred[:, 0] > 0) == (T[:, 0] > 0.5)).mean()
    log(f"training accuracy of the direction choice: {acc * 100:.1f}%")
    save_model(args.out, mlp.params_bf16(), mean, scale, flags=flags)
    # held-out check on fresh graphs (not in the training set)
    held = training_specs(10, seed=99, powerlaw=True)
    eng = BFS("-max_memory=262144")
    summary = []
    for spec in held:
        build(eng, spec)
        model_for_graph(eng, args.out)
        res = compare(eng, pick_sources(eng, 3, seed=spec["seed"]))
        f = sum(r["fixed_ms"] for r in res)
        l = sum(r["learned_ms"] for r in res)
        summary.append({"spec": spec, "fixed_ms": f, "learned_ms": l, "speedup": f / l})
        log(f"held-out {spec['kind']}: fixed {f:.3f} ms, learned {l:.3f} ms, x{f / l:.3f}")
    print(json.dumps({"records": len(X), "train_accuracy": float(acc), "losses": losses,
                      "held_out": summary}))


if __name__ == "__main__":
    main()
