"""Pins the oracle (oracle/oracle.c) before anything is compared against it.

Each oracle function is checked against something independent of it:
scipy's shortest paths (committed fixtures), closed forms, brute force,
and the reference's own page generator compiled from source.
"""
import hashlib
import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

from tests import _oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
U = O.UNREACHED


def load_cases():
    z = np.load(os.path.join(GOLDEN, "bfs_cases.npz"), allow_pickle=False)
    for name in z["names"]:
        yield (str(name), z[f"{name}__offsets"], z[f"{name}__ngh"], int(z[f"{name}__src"][0]),
               z[f"{name}__levels"])


@pytest.mark.parametrize("case", list(load_cases()), ids=lambda c: c[0])
def test_oracle_bfs_matches_scipy_fixtures(case):
    name, off, ngh, src, want = case
    got = O.bfs(off, ngh, src)
    np.testing.assert_array_equal(got, want)


def test_oracle_bfs_closed_forms():
    # path 0->1->...->n-1: level i
    n = 1000
    off, ngh = O.csr_from_edges(n, np.arange(n - 1), np.arange(1, n))
    np.testing.assert_array_equal(O.bfs(off, ngh, 0), np.arange(n, dtype=np.uint32))
    # binary tree: floor(log2(i+1))
    n = 1023
    ch = np.arange(1, n)
    off, ngh = O.csr_from_edges(n, (ch - 1) // 2, ch)
    want = np.floor(np.log2(np.arange(n) + 1)).astype(np.uint32)
    np.testing.assert_array_equal(O.bfs(off, ngh, 0), want)
    # directed cycle from s: (v - s) mod n
    n, s = 77, 30
    off, ngh = O.csr_from_edges(n, np.arange(n), (np.arange(n) + 1) % n)
    np.testing.assert_array_equal(O.bfs(off, ngh, s), ((np.arange(n) - s) % n).astype(np.uint32))
    # grid with both directions: Manhattan distance
    k = 12
    e = []
    for r in range(k):
        for c in range(k):
            v = r * k + c
            if c + 1 < k:
                e += [(v, v + 1), (v + 1, v)]
            if r + 1 < k:
                e += [(v, v + k), (v + k, v)]
    e = np.array(e)
    off, ngh = O.csr_from_edges(k * k, e[:, 0], e[:, 1])
    rr, cc = np.divmod(np.arange(k * k), k)
    np.testing.assert_array_equal(O.bfs(off, ngh, 0), (rr + cc).astype(np.uint32))


def brute_force_levels(n, src, dst, s):
    """Bellman-Ford style relaxation to a fixpoint (unit weights)."""
    d = [None] * n
    d[s] = 0
    changed = True
    while changed:
        changed = False
        for u, v in zip(src, dst):
            if d[u] is not None and (d[v] is None or d[u] + 1 < d[v]):
                d[v] = d[u] + 1
                changed = True
    return np.array([U if x is None else x for x in d], np.uint32)


def test_oracle_bfs_brute_force_tiny():
    rng = np.random.default_rng(5)
    for trial in range(200):
        n = int(rng.integers(1, 13))
        m = int(rng.integers(0, 3 * n + 1))
        s = rng.integers(0, n, m)
        d = rng.integers(0, n, m)
        off, ngh = O.csr_from_edges(n, s, d)
        src = int(rng.integers(0, n))
        np.testing.assert_array_equal(O.bfs(off, ngh, src), brute_force_levels(n, s, d, src))


def test_oracle_bfs_depth_cap_63():
    # The reference starts from MAXDEPTH = 63 (app/bfs.cu:18): a vertex at
    # true depth >= 63 keeps 63, indistinguishable from "unreached".
    n = 100
    off, ngh = O.csr_from_edges(n, np.arange(n - 1), np.arange(1, n))
    capped = O.bfs(off, ngh, 0, inf=63)
    want = np.minimum(np.arange(n), 63).astype(np.uint32)
    np.testing.assert_array_equal(capped, want)
    exact = O.bfs(off, ngh, 0)
    np.testing.assert_array_equal(np.where(exact <= 62, exact, 63), capped)


def test_oracle_bfs_mt_equals_oracle_bfs():
    """The threaded full-size checker gives oracle_bfs's distances: on the
    scipy fixtures, R-MAT and ER graphs with hub out-lists, random digraphs
    with self-loops and duplicates, for 1..16 threads, and under the cap."""
    for name, off, ngh, src, want in load_cases():
        np.testing.assert_array_equal(O.bfs_mt(off, ngh, src, threads=3), want, err_msg=name)
    graphs = [O.rmat_csr(16, 16, seed=4), O.er_csr(50000, 400000, 9)]
    for off, ngh in graphs:
        deg = np.diff(off.astype(np.int64))
        for src in (int(np.argmax(deg)), int(np.nonzero(deg)[0][3]), 0):
            want = O.bfs(off, ngh, src)
            for t in (1, 2, 7, 16):
                np.testing.assert_array_equal(O.bfs_mt(off, ngh, src, threads=t), want)
    rng = np.random.default_rng(12)
    for trial in range(100):
        n = int(rng.integers(1, 40))
        m = int(rng.integers(0, 4 * n + 1))
        off, ngh = O.csr_from_edges(n, rng.integers(0, n, m), rng.integers(0, n, m))
        src = int(rng.integers(0, n))
        np.testing.assert_array_equal(O.bfs_mt(off, ngh, src, threads=4), O.bfs(off, ngh, src))
    n = 100
    off, ngh = O.csr_from_edges(n, np.arange(n - 1), np.arange(1, n))
    np.testing.assert_array_equal(O.bfs_mt(off, ngh, 0, inf=63, threads=4),
                                  O.bfs(off, ngh, 0, inf=63))


# ------------------------------------------------------------ page layout

def ref_tool_available():
    return os.path.exists(O.REF_TOOL)


def _write_inputs(base, off, ngh):
    off.astype(np.uint32).tofile(base + "_csr.idx")
    ngh.astype(np.uint32).tofile(base + "_csr.ngh")
    np.diff(off).astype(np.uint32).tofile(base + "_csr.deg")
    np.zeros(len(ngh), np.uint32).tofile(base + "_csr.wgh")


def test_oracle_pages_match_golden_hashes():
    """Pages built by the oracle hash equal to the reference tool's files."""
    with open(os.path.join(GOLDEN, "pages.json")) as f:
        gold = json.load(f)
    for name, g in gold.items():
        spec = g["spec"]
        if spec["kind"] == "rmat":
            off, ngh = O.rmat_csr(spec["scale"], spec["ef"], seed=spec["seed"])
        else:
            off, ngh = O.er_csr(spec["n"], spec["m"], spec["seed"])
        P = g["pagesize_mb"] << 20
        pages = O.page_cuts(off, P)
        assert pages.tolist() == g["pages"], name
        assert hashlib.sha256(pages.tobytes()).hexdigest() == g["sha256"]["page.info"]
        for i, (left, right, _, _) in enumerate(pages):
            buf = O.fill_page(off, ngh, int(left), int(right), P)
            assert hashlib.sha256(buf.tobytes()).hexdigest() == g["sha256"][f"page{i}"], (name, i)
        info = np.zeros(4, np.uint32)
        info[0] = P
        info[1] = len(off) - 1
        info.view(np.uint64)[1] = len(ngh)
        assert hashlib.sha256(info.tobytes()).hexdigest() == g["sha256"]["graph.info"]


@pytest.mark.skipif(not ref_tool_available(), reason="reference tool not built here")
def test_oracle_page_cuts_match_reference_tool_edge_cases():
    # skewed degrees around the page boundary, tiny pages (1 Mi words)
    rng = np.random.default_rng(3)
    for trial in range(4):
        n = int(rng.integers(5000, 40000))
        deg = rng.zipf(1.7, n).clip(0, 200000)
        deg[rng.integers(0, n, 3)] = int(rng.integers(100000, 900000))
        m = int(deg.sum())
        src = np.repeat(np.arange(n), deg)
        dst = rng.integers(0, n, m)
        off, ngh = O.csr_from_edges(n, src, dst)
        with tempfile.TemporaryDirectory() as td:
            base = os.path.join(td, "g")
            _write_inputs(base, off, ngh)
            subprocess.run([O.REF_TOOL, base, "1", "u"], check=True, stdout=subprocess.DEVNULL)
            ref_pages = np.fromfile(base + "page.info", np.uint32).reshape(-1, 4)
            mine = O.page_cuts(off, 1 << 20)
            np.testing.assert_array_equal(mine, ref_pages)
            for i, (l, r, _, _) in enumerate(mine):
                ref = np.fromfile(base + f"page{i}", np.uint32)
                np.testing.assert_array_equal(O.fill_page(off, ngh, int(l), int(r), 1 << 20), ref)


# ---------------------------------------------------------- page counts

def test_oracle_page_counts_closed_form():
    rng = np.random.default_rng(9)
    off, ngh = O.er_csr(50000, 400000, 4)
    deg = np.diff(off).astype(np.uint32)
    pages = O.page_cuts(off, 1 << 16)
    stat = (rng.random(50000) < 0.1).astype(np.uint8)
    nodes, datas, tot = O.page_counts(pages, stat, deg)
    for p, (l, r, _, _) in enumerate(pages):
        s = stat[l:r + 1].astype(np.int64)
        assert nodes[p] == s.sum()
        assert datas[p] == (s * (deg[l:r + 1].astype(np.int64) + 3)).sum()
    assert tot[0] == stat.sum() and tot[1] == datas.sum()
    _, dw, _ = O.page_counts(pages, stat, deg, weighted=True)
    for p, (l, r, _, _) in enumerate(pages):
        s = stat[l:r + 1].astype(np.int64)
        assert dw[p] == (s * (2 * deg[l:r + 1].astype(np.int64) + 3)).sum()


def test_oracle_compress_closed_form():
    rng = np.random.default_rng(2)
    off, ngh = O.er_csr(5000, 40000, 8)
    P = 1 << 17
    pages = O.page_cuts(off, P)
    l, r = int(pages[0][0]), int(pages[0][1])
    page = O.fill_page(off, ngh, l, r, P)
    stat = (rng.random(5000) < 0.2).astype(np.uint8)
    c, dst = O.compress_page(page, stat, l, r - l + 1, P)
    act = np.nonzero(stat[l:r + 1])[0] + l
    assert c == len(act)
    e = dst[3 * c:]
    pos = 0
    for k, v in enumerate(act):
        d = int(off[v + 1] - off[v])
        assert dst[3 * k] == v and dst[3 * k + 1] == pos and dst[3 * k + 2] == d
        np.testing.assert_array_equal(e[pos:pos + d], ngh[off[v]:off[v + 1]])
        pos += d


# --------------------------------------------------------- chunk masks

def bf_max_zero(mask):
    best = run = 0
    for i in range(32):
        if not (mask >> i) & 1:
            run += 1
            best = max(best, run)
        else:
            run = 0
    return best


def bf_fit(mask, req):
    """Best fit of src/tools.h:29-64 stated directly: among the zero runs of
    length >= req, the first run of exactly req if there is one (the scan
    stops there), else the LAST of the shortest runs (ties replace, since
    the comparison is <=)."""
    runs = []
    i = 0
    while i < 32:
        if not (mask >> i) & 1:
            j = i
            while j < 32 and not (mask >> j) & 1:
                j += 1
            runs.append((i, j - i))
            i = j
        else:
            i += 1
    qual = [(st, ln) for st, ln in runs if ln >= req]
    if not qual:
        return -1
    exact = [st for st, ln in qual if ln == req]
    if exact:
        return exact[0]
    shortest = min(ln for _, ln in qual)
    return [st for st, ln in qual if ln == shortest][-1]


def to_i32(m):
    return int(np.int64(m).astype(np.uint32).view(np.int32))


def test_oracle_chunk_masks_brute_force():
    L = O.lib()
    rng = np.random.default_rng(1)
    masks = list(range(1 << 16)) + [int(x) for x in rng.integers(0, 1 << 32, 20000, dtype=np.int64)]
    masks += [0xFFFFFFFF, 0x80000000, 0x7FFFFFFF, 0x0000FFFF, 0xFFFF0000]
    for m in masks:
        mi = to_i32(m)
        assert L.oracle_max_zero_run(mi) == bf_max_zero(m)
    for m in masks[::7]:
        mi = to_i32(m)
        for req in (1, 2, 3, 5, 8, 16, 31, 32):
            assert L.oracle_fit_zero_run(mi, req) == bf_fit(m, req), (hex(m), req)


def test_oracle_set_clear_runs():
    L = O.lib()
    for off in range(32):
        for num in range(0, 33 - off):
            want = ((1 << num) - 1) << off
            assert (L.oracle_set_run(0, off, num) & 0xFFFFFFFF) == want
            assert (L.oracle_clear_run(-1, off, num) & 0xFFFFFFFF) == (0xFFFFFFFF & ~want)


def test_oracle_pool_size_closed_form():
    L = O.lib()
    P = 10 << 20
    n, pages = 1 << 20, 2
    budget = 4096 * 2 ** 20 - 9 * n - 24 * pages
    assert L.oracle_pool_size(4096, n, pages, P) == min(pages, budget // (4 * P))
    assert L.oracle_pool_size(64, 1 << 26, 224, P) == -1
    assert L.oracle_pool_size(16000, 105_000_000, 383, P) == (16000 * 2 ** 20 - 9 * 105_000_000 -
                                                              24 * 383) // (4 * P)


# ----------------------------------------------------------- generators

def test_splitmix64_definition():
    # SplitMix64 with state 0: the first output is the finaliser applied to
    # the golden gamma 0x9E3779B97F4A7C15 (Steele et al. 2014).
    L = O.lib()
    z = 0x9E3779B97F4A7C15
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2 ** 64 - 1)
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2 ** 64 - 1)
    z ^= z >> 31
    assert L.oracle_splitmix64(0, 0) == z
    assert z == 0xE220A8397B1DCDAF  # published first output of splitmix64 seeded with 0


def test_permutation_is_bijection():
    L = O.lib()
    for bits, n in ((10, 1024), (12, 3000), (7, 100), (1, 2)):
        img = [L.oracle_perm(99, bits, x, n) for x in range(n)]
        assert sorted(img) == list(range(n))


def test_rmat_quadrant_frequencies():
    L = O.lib()
    a, b, c = 0.57, 0.19, 0.19
    A, AB, ABC = O.rmat_thresholds(a, b, c)
    import ctypes
    u = ctypes.c_uint32()
    v = ctypes.c_uint32()
    cnt = np.zeros(4)
    for e in range(20000):
        L.oracle_rmat_edge(5, e, 1, A, AB, ABC, ctypes.byref(u), ctypes.byref(v))
        cnt[2 * u.value + v.value] += 1
    freq = cnt / cnt.sum()
    np.testing.assert_allclose(freq, [a, b, c, 1 - a - b - c], atol=0.015)
    # top bit of a scale-S graph follows the same split
    cnt[:] = 0
    for e in range(20000):
        L.oracle_rmat_edge(5, e, 12, A, AB, ABC, ctypes.byref(u), ctypes.byref(v))
        cnt[2 * (u.value >> 11) + (v.value >> 11)] += 1
    np.testing.assert_allclose(cnt / cnt.sum(), [a, b, c, 1 - a - b - c], atol=0.015)


def test_rmat_csr_is_simple_and_sorted():
    off, ngh = O.rmat_csr(12, 8, seed=3)
    n = len(off) - 1
    assert off[0] == 0 and off[-1] == len(ngh)
    for v in range(n):
        seg = ngh[off[v]:off[v + 1]].astype(np.int64)
        assert np.all(np.diff(seg) > 0)  # sorted, no duplicates
        assert not np.any(seg == v)      # no self-loops


# ------------------------------------------------------------------- MLP

def numpy_mlp(p, x):
    W1 = p[:1024].reshape(64, 16); b1 = p[1024:1088]
    W2 = p[1088:5184].reshape(64, 64); b2 = p[5184:5248]
    W3 = p[5248:5376].reshape(2, 64); b3 = p[5376:5378]
    h1 = np.maximum(x @ W1.T + b1, 0)
    h2 = np.maximum(h1 @ W2.T + b2, 0)
    return h2 @ W3.T + b3


def test_oracle_mlp_forward_matches_matmul_form():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.3, 5378).astype(np.float32)
    x = rng.normal(0, 1, (37, 16)).astype(np.float32)
    y, _, _ = O.mlp_forward(p, x)
    np.testing.assert_allclose(y, numpy_mlp(p.astype(np.float64), x.astype(np.float64)),
                               rtol=1e-4, atol=1e-4)


def test_oracle_mlp_gradient_finite_differences():
    rng = np.random.default_rng(1)
    p = rng.normal(0, 0.3, 5378).astype(np.float64)
    x = rng.normal(0, 1, (8, 16))
    t = np.stack([rng.uniform(0, 1, 8), rng.uniform(-1, 1, 8)], 1)

    def loss(pp):
        y = numpy_mlp(pp, x)
        z0, z1 = y[:, 0], y[:, 1]
        return np.mean(np.logaddexp(0, z0) - t[:, 0] * z0 + 0.5 * (z1 - t[:, 1]) ** 2)

    l32, g = O.mlp_loss_grad(p.astype(np.float32), x.astype(np.float32), t.astype(np.float32), 0.5)
    assert abs(l32 - loss(p)) < 1e-3 * max(1.0, abs(loss(p)))
    idx = rng.choice(5378, 60, replace=False)
    idx = np.concatenate([idx, [0, 1023, 1024, 1100, 5184, 5300, 5376, 5377]])
    for i in idx:
        e = np.zeros_like(p)
        e[i] = 1e-4
        fd = (loss(p + e) - loss(p - e)) / 2e-4
        assert abs(fd - g[i]) <= 2e-3 * max(1.0, abs(fd)), (i, fd, g[i])


# ------------------------------------------------------------ power law

def test_oracle_zipf_degree_tail_and_mean():
    """Pareto tail P(d >= k) = (x_min/k)^(alpha-1) for k below the cap, and
    the capped mean E[min(floor(X), cap)] in closed form."""
    n, x_min, alpha, cap = 200000, 1.2, 1.9, 1000
    d = O.zipf_degrees(n, x_min, alpha, cap, seed=3).astype(np.float64)
    a = alpha - 1
    for k in (2, 5, 20, 100):
        p = (d >= k).mean()
        want = (x_min / k) ** a
        assert abs(p - want) < 4 * np.sqrt(want * (1 - want) / n) + 1e-3, (k, p, want)
    assert d.min() >= 1 and d.max() == cap
    # E[min(floor X, C)] = sum_{k=1..C} P(X >= k) with P(X >= k) = min(1, (x_min/k)^a)
    ks = np.arange(1, cap + 1, dtype=np.float64)
    mean = np.minimum(1.0, (x_min / ks) ** a).sum()
    assert abs(d.mean() - mean) < 0.03 * mean, (d.mean(), mean)


def test_oracle_zipf_csr_is_the_degree_stream():
    n = 3000
    off, ngh = O.zipf_csr(n, 1.2, 1.9, 500, seed=5)
    deg = O.zipf_degrees(n, 1.2, 1.9, 500, seed=5)
    got = np.diff(off.astype(np.int64))
    # dedup and self loops only ever remove edges
    assert np.all(got <= deg) and got.sum() > 0.97 * deg.sum()
    for v in range(0, n, 97):
        row = ngh[off[v]:off[v + 1]]
        assert np.all(np.diff(row.astype(np.int64)) > 0) and v not in row


def weighted_csr(n, src, dst, wgh):
    """CSR with weights carried in the same stable order as csr_from_edges."""
    off, ngh = O.csr_from_edges(n, src, dst)
    order = np.argsort(np.asarray(src, np.int64), kind="stable")
    return off, ngh, np.asarray(wgh, np.uint32)[order]


def test_oracle_wcc_matches_scipy_components():
    # symmetric graphs: the least vertex id of each connected component
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import connected_components
    rng = np.random.default_rng(11)
    for n, m in ((1, 0), (50, 20), (400, 300), (2000, 1500), (3000, 9000)):
        s = rng.integers(0, n, m)
        d = rng.integers(0, n, m)
        ss, dd = np.concatenate([s, d]), np.concatenate([d, s])
        off, ngh = O.csr_from_edges(n, ss, dd)
        ncomp, lab = connected_components(
            csr_matrix((np.ones(len(ss)), (ss, dd)), shape=(n, n)), directed=False)
        least = np.full(ncomp, n, np.int64)
        np.minimum.at(least, lab, np.arange(n))
        np.testing.assert_array_equal(O.wcc(off, ngh), least[lab].astype(np.uint32))


def test_oracle_wcc_brute_force_directed():
    # directed graphs: the least id among the vertices that reach v (v included)
    rng = np.random.default_rng(12)
    for trial in range(150):
        n = int(rng.integers(1, 12))
        m = int(rng.integers(0, 3 * n + 1))
        s = rng.integers(0, n, m)
        d = rng.integers(0, n, m)
        reach = np.eye(n, dtype=bool)
        reach[s, d] = True
        for k in range(n):  # Warshall closure
            reach |= reach[:, [k]] & reach[[k], :]
        want = np.array([np.flatnonzero(reach[:, v]).min() for v in range(n)], np.uint32)
        off, ngh = O.csr_from_edges(n, s, d)
        np.testing.assert_array_equal(O.wcc(off, ngh), want)


def test_oracle_sssp_matches_scipy_dijkstra():
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import dijkstra
    rng = np.random.default_rng(13)
    for n, m, wmax in ((30, 60, 5), (500, 2500, 100), (3000, 20000, 1000), (2000, 800, 7)):
        s = rng.integers(0, n, m)
        d = rng.integers(0, n, m)
        w = rng.integers(1, wmax + 1, m)
        # duplicate edges: scipy would sum them, keep the lightest instead
        key = s * n + d
        order = np.lexsort((w, key))
        first = np.ones(m, bool)
        first[1:] = key[order][1:] != key[order][:-1]
        keep = order[first]
        mat = csr_matrix((w[keep].astype(np.float64), (s[keep], d[keep])), shape=(n, n))
        off, ngh, wg = weighted_csr(n, s, d, w)
        for src in (0, int(rng.integers(0, n))):
            dist = dijkstra(mat, directed=True, indices=src)
            want = np.where(np.isinf(dist), O.SSSP_INF, dist).astype(np.uint32)
            np.testing.assert_array_equal(O.sssp(off, ngh, wg, src), want)


def test_oracle_sssp_brute_force_zero_weights():
    # Bellman-Ford on tiny graphs, zero weights and self-loops included
    rng = np.random.default_rng(14)
    for trial in range(200):
        n = int(rng.integers(1, 12))
        m = int(rng.integers(0, 3 * n + 1))
        s = rng.integers(0, n, m)
        d = rng.integers(0, n, m)
        w = rng.integers(0, 4, m)
        src = int(rng.integers(0, n))
        dist = [None] * n
        dist[src] = 0
        for _ in range(n):
            for u, v, x in zip(s, d, w):
                if dist[u] is not None and (dist[v] is None or dist[u] + x < dist[v]):
                    dist[v] = dist[u] + int(x)
        want = np.array([O.SSSP_INF if x is None else x for x in dist], np.uint32)
        off, ngh, wg = weighted_csr(n, s, d, w)
        np.testing.assert_array_equal(O.sssp(off, ngh, wg, src), want)


def test_oracle_sssp_unit_weights_is_bfs():
    for case in load_cases():
        name, off, ngh, src, levels = case
        got = O.sssp(off, ngh, np.ones(len(ngh), np.uint32), src)
        want = np.where(levels == U, O.SSSP_INF, levels).astype(np.uint32)
        np.testing.assert_array_equal(got, want, err_msg=name)
