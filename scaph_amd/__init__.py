"""scaph_amd — MI355X-native BFS hot path of the Scaph out-of-core graph engine.

Python mirror of the reference's App interface (src/app.cuh:120-135,
app/bfs.cu:39-117) over the C-ABI engine in ``libscaph_host.so``:

    eng = BFS("-max_memory=65536", "-schedule=fixed")
    eng.load_csr(offsets, ngh)          # or eng.load() for -graphfile / -gen
    stats = eng.run(src)                # one traversal (t_BFS, n_c, m_c)
    levels = eng.levels()               # uint32, 0xFFFFFFFF = unreached
    eng.check(src)                      # CPU BFS comparison (bfs::check)
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import UNREACHED, LibraryMissing, ScaphError  # noqa: F401

__all__ = ["BFS", "Batch", "generate_rmat", "generate_er", "page_cuts", "plan_pool", "write_graph",
           "UNREACHED",
           "LibraryMissing", "ScaphError", "device_props"]


def _err():
    return ctypes.create_string_buffer(1024)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class BFS:
    """The BFS app: construct from gflags-style flags, load a graph, run."""

    def __init__(self, *flags):
        lib = _lib.host()
        argv = [b"bfs"] + [f.encode() for f in flags]
        arr = (ctypes.c_char_p * len(argv))(*argv)
        h = ctypes.c_void_p()
        e = _err()
        _lib.check(lib.scaph_engine_create(len(argv), arr, ctypes.byref(h), e, 1024), e.value.decode())
        self._h = h
        self._lib = lib

    def close(self):
        if getattr(self, "_h", None):
            self._lib.scaph_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- load_graph ------------------------------------------------------
    def load(self):
        e = _err()
        _lib.check(self._lib.scaph_engine_load(self._h, e, 1024), e.value.decode())

    def load_csr(self, offsets, ngh, keep_host=True):
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        ngh = np.ascontiguousarray(ngh, dtype=np.uint32)
        n = len(offsets) - 1
        if ngh.size == 0:
            ngh = np.zeros(1, np.uint32)
        e = _err()
        _lib.check(self._lib.scaph_engine_load_csr(self._h, n, _ptr(offsets), _ptr(ngh),
                                                  int(keep_host), e, 1024), e.value.decode())

    def gen_rmat_device(self, scale, edge_factor, a=0.57, b=0.19, c=0.19, seed=1, keep_host=False):
        e = _err()
        _lib.check(self._lib.scaph_engine_gen_rmat_device(self._h, scale, edge_factor, a, b, c,
                                                         seed, int(keep_host), e, 1024),
                   e.value.decode())

    def gen_er_device(self, n, m, seed=1, keep_host=False):
        e = _err()
        _lib.check(self._lib.scaph_engine_gen_er_device(self._h, n, m, seed, int(keep_host), e,
                                                       1024), e.value.decode())

    def gen_zipf(self, n, x_min=1.2, alpha=1.9, cap=1000000, seed=3):
        """Power-law digraph (Pareto out-degrees, uniform destinations)."""
        e = _err()
        _lib.check(self._lib.scaph_engine_gen_zipf(self._h, n, x_min, alpha, cap, seed, e, 1024),
                   e.value.decode())

    def info(self):
        n = ctypes.c_uint32()
        m = ctypes.c_uint64()
        np_ = ctypes.c_uint32()
        res = ctypes.c_int()
        pull = ctypes.c_int()
        _lib.check(self._lib.scaph_engine_info(self._h, ctypes.byref(n), ctypes.byref(m),
                                               ctypes.byref(np_), ctypes.byref(res),
                                               ctypes.byref(pull)))
        return {"n": n.value, "m": m.value, "pages": np_.value, "resident": bool(res.value),
                "pull": bool(pull.value)}

    def pool(self):
        """Page slots of the device pool (every page when resident) and the
        device bytes the engine holds."""
        slots = ctypes.c_int64()
        nbytes = ctypes.c_uint64()
        _lib.check(self._lib.scaph_engine_pool(self._h, ctypes.byref(slots), ctypes.byref(nbytes)))
        return {"slots": slots.value, "device_bytes": nbytes.value}

    # -- run -------------------------------------------------------------
    def run(self, src):
        st = _lib.BfsStats()
        e = _err()
        _lib.check(self._lib.scaph_engine_bfs(self._h, src, ctypes.byref(st), e, 1024),
                   e.value.decode())
        return {k: getattr(st, k) for k, _ in _lib.BfsStats._fields_}

    # -- the reference's other App plugins (app/wcc.cu, app/sssp.cu) --------
    def set_weights(self, wgh):
        """Edge weights in CSR edge order (the _csr.wgh file), for sssp()."""
        w = np.ascontiguousarray(wgh, dtype=np.uint32)
        e = _err()
        _lib.check(self._lib.scaph_engine_set_weights(self._h, _ptr(w), w.size, e, 1024),
                   e.value.decode())

    def wcc(self):
        """Least vertex id reaching each vertex (read with levels())."""
        st = _lib.BfsStats()
        e = _err()
        _lib.check(self._lib.scaph_engine_wcc(self._h, ctypes.byref(st), e, 1024),
                   e.value.decode())
        return {k: getattr(st, k) for k, _ in _lib.BfsStats._fields_}

    def sssp(self, src):
        """Shortest-path lengths over the weights (0x7FFFFFFF unreached; levels())."""
        st = _lib.BfsStats()
        e = _err()
        _lib.check(self._lib.scaph_engine_sssp(self._h, src, ctypes.byref(st), e, 1024),
                   e.value.decode())
        return {k: getattr(st, k) for k, _ in _lib.BfsStats._fields_}

    def set_schedule(self, name, alpha=0.0, beta=0.0):
        e = _err()
        _lib.check(self._lib.scaph_engine_set_schedule(self._h, name.encode(), alpha, beta, e,
                                                       1024), e.value.decode())

    def set_page_levels(self, levels):
        """-page_predict=exact: the levels of an earlier BFS from the next
        source, so the paged scheduler knows each page's future activity."""
        lv = np.ascontiguousarray(levels, dtype=np.uint32)
        e = _err()
        _lib.check(self._lib.scaph_engine_set_page_levels(self._h, _ptr(lv), lv.size, e, 1024),
                   e.value.decode())

    def set_model(self, params_bf16, mean, scale, prefetch_gain=0.0):
        p = np.ascontiguousarray(params_bf16, dtype=np.uint16)
        m = np.ascontiguousarray(mean, dtype=np.float32)
        s = np.ascontiguousarray(scale, dtype=np.float32)
        e = _err()
        _lib.check(self._lib.scaph_engine_set_model(self._h, _ptr(p), _ptr(m), _ptr(s),
                                                    prefetch_gain, e, 1024), e.value.decode())

    def records(self):
        cnt = ctypes.c_uint32()
        self._lib.scaph_engine_level_records(self._h, None, 0, ctypes.byref(cnt))
        arr = (_lib.LevelRecord * max(cnt.value, 1))()
        self._lib.scaph_engine_level_records(self._h, arr, cnt.value, ctypes.byref(cnt))
        out = []
        for r in arr[:cnt.value]:
            d = {k: getattr(r, k) for k, _ in _lib.LevelRecord._fields_
                 if k not in ("page_hist", "pad")}
            d["page_hist"] = list(r.page_hist)
            out.append(d)
        return out

    def explore(self, src, reps=2):
        """Training traces: per level, both directions timed from the same state."""
        arr = (_lib.ExploreRecord * 4096)()
        cnt = ctypes.c_uint32()
        e = _err()
        _lib.check(self._lib.scaph_engine_explore(self._h, src, reps, arr, 4096, ctypes.byref(cnt),
                                                  e, 1024), e.value.decode())
        out = []
        for r in arr[:min(cnt.value, 4096)]:
            out.append({"level": r.level, "n_f": r.n_f, "m_f": r.m_f, "n_next": r.n_next,
                        "t_push_ms": r.t_push_ms, "t_pull_ms": r.t_pull_ms,
                        "pages_active_next": r.pages_active_next, "x": list(r.x)})
        return out

    # -- multi-source (batched) BFS ------------------------------------------
    def msbfs(self, sources, tune=None):
        """Up to 256 BFS queries in one traversal; returns the run's stats."""
        src = np.ascontiguousarray(sources, dtype=np.uint32)
        st = _lib.MsStats()
        t = None
        if tune is not None:
            t = _lib.MsTune(tune.get("push_div", 64), tune.get("push_mdiv", 16),
                            tune.get("pull_block", 256), tune.get("use_graph", 0),
                            tune.get("segment", 8), tune.get("pull_grid", 0),
                            tune.get("persistent", 0))
        e = _err()
        _lib.check(self._lib.scaph_engine_msbfs(self._h, _ptr(src), len(src),
                                                ctypes.byref(t) if t is not None else None,
                                                ctypes.byref(st), e, 1024), e.value.decode())
        return {"ms": st.ms, "depth": st.depth, "queries": st.queries, "pairs": st.pairs}

    def msbfs_levels(self, query):
        n = self.info()["n"]
        out = np.empty(n, np.uint32)
        e = _err()
        _lib.check(self._lib.scaph_engine_msbfs_levels(self._h, query, _ptr(out), e, 1024),
                   e.value.decode())
        return out

    def msbfs_counters(self):
        arr = (_lib.MsCounters * 256)()
        cnt = ctypes.c_uint32()
        _lib.check(self._lib.scaph_engine_msbfs_counters(self._h, arr, 256, ctypes.byref(cnt)))
        return [{k: getattr(r, k) for k in ("n_active", "bits", "m_active", "q_len", "heavy_len",
                                             "overflow")} for r in arr[:min(cnt.value, 256)]]

    def msbfs_autotune(self, sources):
        src = np.ascontiguousarray(sources, dtype=np.uint32)
        best = _lib.MsTune()
        bms = ctypes.c_double()
        logt = (_lib.MsTune * 128)()
        logm = (ctypes.c_double * 128)()
        cnt = ctypes.c_uint32()
        e = _err()
        _lib.check(self._lib.scaph_engine_msbfs_autotune(self._h, _ptr(src), len(src),
                                                         ctypes.byref(best), ctypes.byref(bms),
                                                         logt, logm, 128, ctypes.byref(cnt), e,
                                                         1024), e.value.decode())
        as_d = lambda t: {k: getattr(t, k) for k, _ in _lib.MsTune._fields_}  # noqa: E731
        return as_d(best), bms.value, [(as_d(logt[i]), logm[i]) for i in range(min(cnt.value, 128))]

    def levels(self):
        n = self.info()["n"]
        out = np.empty(n, np.uint32)
        _lib.check(self._lib.scaph_engine_levels(self._h, _ptr(out)))
        return out

    def out_degrees(self):
        n = self.info()["n"]
        out = np.empty(n, np.uint32)
        _lib.check(self._lib.scaph_engine_out_degrees(self._h, _ptr(out)))
        return out

    def host_csr(self):
        inf = self.info()
        off = np.empty(inf["n"] + 1, np.uint64)
        ngh = np.empty(max(inf["m"], 1), np.uint32)
        _lib.check(self._lib.scaph_engine_host_csr(self._h, _ptr(off), _ptr(ngh)),
                   "no host CSR kept")
        return off, ngh[:inf["m"]]

    def validate(self, src):
        out = np.zeros(2, np.uint64)
        _lib.check(self._lib.scaph_engine_validate(self._h, src, _ptr(out)))
        return int(out[0]), int(out[1])

    def check(self, src):
        errors = ctypes.c_uint64()
        e = _err()
        _lib.check(self._lib.scaph_engine_check(self._h, src, ctypes.byref(errors), e, 1024),
                   e.value.decode())
        return errors.value

    def graph_desc(self, transposed=False):
        d = _lib.GraphDesc()
        _lib.check(self._lib.scaph_engine_graph_desc(self._h, int(transposed), ctypes.byref(d)))
        return d

    def device_degree(self):
        p = ctypes.c_void_p()
        _lib.check(self._lib.scaph_engine_device_degree(self._h, ctypes.byref(p)))
        return p.value


class Batch:
    """A batch of small R-MAT graphs on the device, all queried by one
    persistent launch (one BFS per graph from its own source)."""

    def __init__(self, count=4096, min_scale=10, max_scale=16, min_ef=8, max_ef=32, seed=1,
                 device=0):
        lib = _lib.host()
        h = ctypes.c_void_p()
        e = _err()
        _lib.check(lib.scaph_batch_create_rmat(count, min_scale, max_scale, min_ef, max_ef, seed,
                                               device, ctypes.byref(h), e, 1024),
                   e.value.decode())
        self._h = h
        self._lib = lib

    def close(self):
        if getattr(self, "_h", None):
            self._lib.scaph_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def info(self):
        c, v, m = ctypes.c_uint32(), ctypes.c_uint64(), ctypes.c_uint64()
        _lib.check(self._lib.scaph_batch_info(self._h, ctypes.byref(c), ctypes.byref(v),
                                              ctypes.byref(m)))
        return {"graphs": c.value, "vertices": v.value, "edges": m.value}

    def graph(self, g):
        r = _lib.BatchEntry()
        _lib.check(self._lib.scaph_batch_graph_info(self._h, g, ctypes.byref(r)))
        return {k: getattr(r, k) for k, _ in _lib.BatchEntry._fields_}

    def run(self, tune=None):
        t = None
        if tune is not None:
            t = _lib.BatchTune(tune.get("block", 256), tune.get("wgs_per_cu", 4),
                               tune.get("use_graph", 0))
        ms = ctypes.c_double()
        e = _err()
        _lib.check(self._lib.scaph_batch_run(self._h, ctypes.byref(t) if t is not None else None,
                                             ctypes.byref(ms), e, 1024), e.value.decode())
        return ms.value

    def levels(self, g):
        out = np.empty(self.graph(g)["n"], np.uint32)
        _lib.check(self._lib.scaph_batch_levels(self._h, g, _ptr(out)))
        return out

    def csr(self, g):
        n = self.graph(g)["n"]
        off = np.empty(n + 1, np.uint64)
        m = ctypes.c_uint64()
        _lib.check(self._lib.scaph_batch_csr(self._h, g, _ptr(off), None, 0, ctypes.byref(m)))
        ngh = np.empty(max(m.value, 1), np.uint32)
        _lib.check(self._lib.scaph_batch_csr(self._h, g, _ptr(off), _ptr(ngh), len(ngh),
                                             ctypes.byref(m)))
        return off, ngh[:m.value]

    def autotune(self):
        best = _lib.BatchTune()
        bms = ctypes.c_double()
        logt = (_lib.BatchTune * 64)()
        logm = (ctypes.c_double * 64)()
        cnt = ctypes.c_uint32()
        e = _err()
        _lib.check(self._lib.scaph_batch_autotune(self._h, ctypes.byref(best), ctypes.byref(bms),
                                                  logt, logm, 64, ctypes.byref(cnt), e, 1024),
                   e.value.decode())
        as_d = lambda t: {k: getattr(t, k) for k, _ in _lib.BatchTune._fields_}  # noqa: E731
        return as_d(best), bms.value, [(as_d(logt[i]), logm[i]) for i in range(min(cnt.value, 64))]


def generate_rmat(scale, edge_factor, a=0.57, b=0.19, c=0.19, seed=1):
    """Host R-MAT generator (same edges as the device builder)."""
    lib = _lib.host()
    n = 1 << scale
    off = np.empty(n + 1, np.uint64)
    cap = edge_factor * n
    ngh = np.empty(max(cap, 1), np.uint32)
    m = ctypes.c_uint64()
    _lib.check(lib.scaph_host_gen_rmat(scale, edge_factor, a, b, c, seed, _ptr(off), _ptr(ngh),
                                       cap, ctypes.byref(m)))
    return off, ngh[:m.value].copy()


def generate_er(n, m, seed=1):
    lib = _lib.host()
    off = np.empty(n + 1, np.uint64)
    ngh = np.empty(max(m, 1), np.uint32)
    E = ctypes.c_uint64()
    _lib.check(lib.scaph_host_gen_er(n, m, seed, _ptr(off), _ptr(ngh), m, ctypes.byref(E)))
    return off, ngh[:E.value].copy()


def page_cuts(offsets, pagesize_words, weighted=False):
    lib = _lib.host()
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    n = len(offsets) - 1
    cnt = ctypes.c_uint32()
    _lib.check(lib.scaph_host_page_cuts(_ptr(offsets), n, pagesize_words, int(weighted), None, 0,
                                        ctypes.byref(cnt)))
    pages = np.zeros((cnt.value, 4), np.uint32)
    _lib.check(lib.scaph_host_page_cuts(_ptr(offsets), n, pagesize_words, int(weighted),
                                        _ptr(pages), cnt.value, ctypes.byref(cnt)))
    return pages


def plan_pool(max_memory_mib, n, num_pages, pagesize_words):
    """Device page slots for -max_memory (src/app.cuh:281-286, 323-324);
    -1 where the reference exits with "vertex data out of gpu memory"."""
    slots = ctypes.c_int64()
    _lib.check(_lib.host().scaph_host_plan_pool(max_memory_mib, n, num_pages, pagesize_words,
                                                ctypes.byref(slots)))
    return slots.value


def write_graph(base, offsets, ngh, pagesize_words=0, idx64=True, write_wgh=False):
    lib = _lib.host()
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    ngh = np.ascontiguousarray(ngh, dtype=np.uint32)
    if ngh.size == 0:
        ngh = np.zeros(1, np.uint32)
    e = _err()
    _lib.check(lib.scaph_host_write_graph(str(base).encode(), len(offsets) - 1, _ptr(offsets),
                                          _ptr(ngh), pagesize_words, int(idx64), int(write_wgh),
                                          e, 1024), e.value.decode())


def device_props(device=0):
    lib = _lib.hip()
    p = _lib.DeviceProps()
    _lib.check(lib.scaph_get_device_props(device, ctypes.byref(p)))
    return {"name": p.name.decode(), "arch": p.arch.decode(), "cu": p.compute_units,
            "total_mem": p.total_mem, "l2": p.l2_bytes, "mem_clock_khz": p.mem_clock_khz,
            "bus_width": p.mem_bus_width}
