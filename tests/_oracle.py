"""ctypes wrapper of oracle/liboracle.so — the CPU checker (test use only)."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")
REF_TOOL = os.path.join(ROOT, "oracle", "_ref", "bin2pagecsr")
UNREACHED = 0xFFFFFFFF

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: run __graft_entry__.build()")
        L = ctypes.CDLL(LIB)
        P = ctypes.c_void_p
        U32, U64, I32 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
        L.oracle_bfs.argtypes = [U32, P, P, U32, U32, P]
        L.oracle_bfs.restype = I32
        L.oracle_bfs_mt.argtypes = [U32, P, P, U32, U32, P, U32]
        L.oracle_bfs_mt.restype = I32
        L.oracle_wcc.argtypes = [U32, P, P, P]
        L.oracle_wcc.restype = I32
        L.oracle_sssp.argtypes = [U32, P, P, P, U32, P]
        L.oracle_sssp.restype = I32
        L.oracle_page_cuts.argtypes = [P, U32, U64, I32, P, U32, ctypes.POINTER(U32)]
        L.oracle_page_cuts.restype = I32
        L.oracle_fill_page.argtypes = [P, P, P, U32, U32, I32, P]
        L.oracle_fill_page.restype = None
        L.oracle_page_counts.argtypes = [P, U32, P, P, I32, P, P, P]
        L.oracle_page_counts.restype = None
        L.oracle_compress_page.argtypes = [P, P, P, U32, U32, U32]
        L.oracle_compress_page.restype = U32
        for n in ("oracle_max_zero_run",):
            getattr(L, n).argtypes = [I32]
            getattr(L, n).restype = I32
        L.oracle_fit_zero_run.argtypes = [I32, I32]
        L.oracle_fit_zero_run.restype = I32
        L.oracle_set_run.argtypes = [I32, I32, I32]
        L.oracle_set_run.restype = I32
        L.oracle_clear_run.argtypes = [I32, I32, I32]
        L.oracle_clear_run.restype = I32
        L.oracle_pool_size.argtypes = [ctypes.c_int64, U32, U32, U64]
        L.oracle_pool_size.restype = ctypes.c_int64
        L.oracle_splitmix64.argtypes = [U64, U64]
        L.oracle_splitmix64.restype = U64
        L.oracle_perm.argtypes = [U64, U32, U64, U64]
        L.oracle_perm.restype = U64
        L.oracle_rmat_edge.argtypes = [U64, U64, U32, U32, U32, U32, ctypes.POINTER(U32),
                                       ctypes.POINTER(U32)]
        L.oracle_rmat_edge.restype = None
        L.oracle_rmat_csr.argtypes = [U32, U32, U32, U32, U32, U64, P, P, U64, ctypes.POINTER(U64)]
        L.oracle_rmat_csr.restype = I32
        L.oracle_er_csr.argtypes = [U32, U64, U64, P, P, U64, ctypes.POINTER(U64)]
        L.oracle_er_csr.restype = I32
        D = ctypes.c_double
        L.oracle_zipf_degree.argtypes = [U64, U32, D, D, U32]
        L.oracle_zipf_degree.restype = U32
        L.oracle_zipf_csr.argtypes = [U32, D, D, U32, U64, P, P, U64, ctypes.POINTER(U64)]
        L.oracle_zipf_csr.restype = I32
        L.oracle_mlp_forward.argtypes = [P, P, U32, P, P, P]
        L.oracle_mlp_forward.restype = None
        L.oracle_mlp_loss_grad.argtypes = [P, P, P, U32, ctypes.c_float, P]
        L.oracle_mlp_loss_grad.restype = ctypes.c_float
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def wcc(offsets, ngh):
    """app/wcc.cu check(): least vertex id reaching each vertex."""
    offsets = np.ascontiguousarray(offsets, np.uint64)
    ngh = np.ascontiguousarray(ngh, np.uint32)
    n = len(offsets) - 1
    out = np.empty(max(n, 1), np.uint32)
    if ngh.size == 0:
        ngh = np.zeros(1, np.uint32)
    rc = lib().oracle_wcc(n, _p(offsets), _p(ngh), _p(out))
    if rc:
        raise ValueError(f"oracle_wcc rc={rc}")
    return out[:n]


SSSP_INF = 0x7FFFFFFF


def sssp(offsets, ngh, wgh, src):
    """Shortest paths over u32 weights, 0x7FFFFFFF for unreached."""
    offsets = np.ascontiguousarray(offsets, np.uint64)
    ngh = np.ascontiguousarray(ngh, np.uint32)
    wgh = np.ascontiguousarray(wgh, np.uint32)
    n = len(offsets) - 1
    out = np.empty(max(n, 1), np.uint32)
    if ngh.size == 0:
        ngh = np.zeros(1, np.uint32)
        wgh = np.zeros(1, np.uint32)
    rc = lib().oracle_sssp(n, _p(offsets), _p(ngh), _p(wgh), src, _p(out))
    if rc:
        raise ValueError(f"oracle_sssp rc={rc}")
    return out[:n]


def bfs(offsets, ngh, src, inf=UNREACHED):
    offsets = np.ascontiguousarray(offsets, np.uint64)
    ngh = np.ascontiguousarray(ngh, np.uint32)
    n = len(offsets) - 1
    out = np.empty(max(n, 1), np.uint32)
    if ngh.size == 0:
        ngh = np.zeros(1, np.uint32)
    rc = lib().oracle_bfs(n, _p(offsets), _p(ngh), src, inf, _p(out))
    if rc:
        raise ValueError(f"oracle_bfs rc={rc}")
    return out[:n]


def bfs_mt(offsets, ngh, src, inf=UNREACHED, threads=None):
    """oracle_bfs's distances, level-synchronous on `threads` threads (for
    full-size graphs; pinned by equality with bfs())."""
    offsets = np.ascontiguousarray(offsets, np.uint64)
    ngh = np.ascontiguousarray(ngh, np.uint32)
    n = len(offsets) - 1
    out = np.empty(max(n, 1), np.uint32)
    if ngh.size == 0:
        ngh = np.zeros(1, np.uint32)
    if threads is None:
        threads = max(1, min(16, len(os.sched_getaffinity(0))))
    rc = lib().oracle_bfs_mt(n, _p(offsets), _p(ngh), src, inf, _p(out), threads)
    if rc:
        raise ValueError(f"oracle_bfs_mt rc={rc}")
    return out[:n]


def page_cuts(offsets, P, weighted=False):
    offsets = np.ascontiguousarray(offsets, np.uint64)
    n = len(offsets) - 1
    cnt = ctypes.c_uint32()
    rc = lib().oracle_page_cuts(_p(offsets), n, P, int(weighted), None, 0, ctypes.byref(cnt))
    if rc:
        raise ValueError(f"oracle_page_cuts rc={rc}")
    pages = np.zeros((max(cnt.value, 1), 4), np.uint32)
    rc = lib().oracle_page_cuts(_p(offsets), n, P, int(weighted), _p(pages), cnt.value,
                                ctypes.byref(cnt))
    return pages[:cnt.value]


def fill_page(offsets, ngh, left, right, P, weighted=False, wgh=None):
    offsets = np.ascontiguousarray(offsets, np.uint64)
    ngh = np.ascontiguousarray(ngh, np.uint32)
    out = np.zeros(P, np.uint32)
    w = None if wgh is None else np.ascontiguousarray(wgh, np.uint32)
    lib().oracle_fill_page(_p(offsets), _p(ngh), None if w is None else _p(w), left, right,
                           int(weighted), _p(out))
    return out


def page_counts(pages, stat, degree, weighted=False):
    pages = np.ascontiguousarray(pages, np.uint32)
    stat = np.ascontiguousarray(stat, np.uint8)
    degree = np.ascontiguousarray(degree, np.uint32)
    np_ = len(pages)
    nodes = np.zeros(max(np_, 1), np.uint32)
    datas = np.zeros(max(np_, 1), np.uint32)
    tot = np.zeros(2, np.uint64)
    lib().oracle_page_counts(_p(pages), np_, _p(stat), _p(degree), int(weighted), _p(nodes),
                             _p(datas), _p(tot))
    return nodes[:np_], datas[:np_], tot


def compress_page(page, stat, left, nodenum, out_words):
    page = np.ascontiguousarray(page, np.uint32)
    stat = np.ascontiguousarray(stat, np.uint8)
    dst = np.zeros(out_words, np.uint32)
    cnt = int(stat[left:left + nodenum].sum())
    c = lib().oracle_compress_page(_p(page), _p(dst), _p(stat), left, nodenum, cnt)
    return c, dst


def rmat_thresholds(a, b, c):
    def thr(x):
        if x <= 0:
            return 0
        t = np.floor(x * 4294967296.0)
        return 0xFFFFFFFF if t >= 4294967295.0 else int(t)
    return thr(a), thr(a + b), thr(a + b + c)


def rmat_csr(scale, ef, a=0.57, b=0.19, c=0.19, seed=1):
    A, AB, ABC = rmat_thresholds(a, b, c)
    n = 1 << scale
    off = np.empty(n + 1, np.uint64)
    cap = ef * n
    ngh = np.empty(max(cap, 1), np.uint32)
    E = ctypes.c_uint64()
    rc = lib().oracle_rmat_csr(scale, ef, A, AB, ABC, seed, _p(off), _p(ngh), cap, ctypes.byref(E))
    if rc:
        raise ValueError(f"oracle_rmat_csr rc={rc}")
    return off, ngh[:E.value].copy()


def er_csr(n, m, seed):
    off = np.empty(n + 1, np.uint64)
    ngh = np.empty(max(m, 1), np.uint32)
    E = ctypes.c_uint64()
    rc = lib().oracle_er_csr(n, m, seed, _p(off), _p(ngh), m, ctypes.byref(E))
    if rc:
        raise ValueError(f"oracle_er_csr rc={rc}")
    return off, ngh[:E.value].copy()


def zipf_degrees(n, x_min, alpha, cap, seed):
    f = lib().oracle_zipf_degree
    return np.array([f(seed, v, x_min, alpha, cap) for v in range(n)], np.uint32)


def zipf_csr(n, x_min, alpha, cap, seed):
    m = int(zipf_degrees(n, x_min, alpha, cap, seed).astype(np.uint64).sum())
    off = np.empty(n + 1, np.uint64)
    ngh = np.empty(max(m, 1), np.uint32)
    E = ctypes.c_uint64()
    rc = lib().oracle_zipf_csr(n, x_min, alpha, cap, seed, _p(off), _p(ngh), m, ctypes.byref(E))
    if rc:
        raise ValueError(f"oracle_zipf_csr rc={rc}")
    return off, ngh[:E.value].copy()


def mlp_forward(params, x):
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    b = x.shape[0]
    y = np.zeros((b, 2), np.float32)
    h1 = np.zeros((b, 64), np.float32)
    h2 = np.zeros((b, 64), np.float32)
    lib().oracle_mlp_forward(_p(params), _p(x), b, _p(y), _p(h1), _p(h2))
    return y, h1, h2


def mlp_loss_grad(params, x, t, w1):
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    t = np.ascontiguousarray(t, np.float32)
    g = np.zeros_like(params)
    loss = lib().oracle_mlp_loss_grad(_p(params), _p(x), _p(t), x.shape[0], w1, _p(g))
    return float(loss), g


def csr_from_edges(n, src, dst):
    """Plain CSR (keeps duplicates and self-loops, stable order)."""
    src = np.asarray(src, np.int64)
    dst = np.asarray(dst, np.int64)
    order = np.argsort(src, kind="stable")
    ngh = dst[order].astype(np.uint32)
    cnt = np.bincount(src, minlength=n)
    off = np.zeros(n + 1, np.uint64)
    off[1:] = np.cumsum(cnt)
    return off, ngh
