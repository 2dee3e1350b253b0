"""Writes the committed golden fixtures (run in the build container only).

bfs_cases.npz   small directed graphs (with duplicate edges, self-loops,
                unreachable parts, a path longer than the reference's depth
                cap of 63) and their BFS levels computed by
                scipy.sparse.csgraph.shortest_path(unweighted=True) — an
                implementation independent of the oracle and the product.
pages.json      SHA-256 of the files the reference's own page generator
                (tools/bin2pagecsr.cpp, compiled from source into
                oracle/_ref/bin2pagecsr) writes for small synthetic graphs,
                so the page layout stays pinned on machines without the
                reference tree.

Usage: python tests/golden/make_golden.py   (needs oracle/_ref built)
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import shortest_path

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from tests import _oracle  # noqa: E402

UNREACHED = 0xFFFFFFFF


def csr(n, src, dst):
    return _oracle.csr_from_edges(n, src, dst)


def scipy_levels(off, ngh, s):
    n = len(off) - 1
    data = np.ones(len(ngh), np.float64)
    A = sp.csr_matrix((data, ngh.astype(np.int64), off.astype(np.int64)), shape=(n, n))
    d = shortest_path(A, directed=True, unweighted=True, indices=s)
    out = np.full(n, UNREACHED, np.uint32)
    fin = np.isfinite(d)
    out[fin] = d[fin].astype(np.uint32)
    return out


def cases():
    rng = np.random.default_rng(20240601)
    out = []
    # random directed graph with duplicates and self-loops
    n = 300
    s = rng.integers(0, n, 2400)
    d = rng.integers(0, n, 2400)
    out.append(("random_dups", n, s, d, 0))
    # sparse random graph with several components
    n = 500
    s = rng.integers(0, n, 600)
    d = rng.integers(0, n, 600)
    out.append(("random_sparse", n, s, d, int(s[0])))
    # directed path of 100 (depth 99 > cap 63)
    n = 100
    out.append(("path100", n, np.arange(n - 1), np.arange(1, n), 0))
    # star out of vertex 7
    n = 64
    leaves = np.array([v for v in range(n) if v != 7])
    out.append(("star", n, np.full(len(leaves), 7), leaves, 7))
    # 2-D grid 16x16, edges both ways
    k = 16
    e = []
    for r in range(k):
        for c in range(k):
            v = r * k + c
            if c + 1 < k:
                e += [(v, v + 1), (v + 1, v)]
            if r + 1 < k:
                e += [(v, v + k), (v + k, v)]
    e = np.array(e)
    out.append(("grid16", k * k, e[:, 0], e[:, 1], 0))
    # complete binary tree of 255 nodes, parent -> child
    n = 255
    ch = np.arange(1, n)
    out.append(("bintree", n, (ch - 1) // 2, ch, 0))
    # directed cycle
    n = 50
    out.append(("cycle50", n, np.arange(n), (np.arange(n) + 1) % n, 13))
    # isolated source
    n = 20
    out.append(("isolated_src", n, np.array([1, 2, 3]), np.array([2, 3, 4]), 0))
    # single vertex with a self loop
    out.append(("single_loop", 1, np.array([0]), np.array([0]), 0))
    # hub: vertex 0 points at everyone, the rest form a sparse graph
    n = 2000
    s = np.concatenate([np.zeros(n - 1, np.int64), rng.integers(1, n, 3000)])
    d = np.concatenate([np.arange(1, n), rng.integers(1, n, 3000)])
    out.append(("hub", n, s, d, int(rng.integers(1, n))))
    return out


def write_bfs_cases():
    arrays = {}
    names = []
    for name, n, s, d, src in cases():
        off, ngh = csr(n, s, d)
        lv = scipy_levels(off, ngh, src)
        arrays[f"{name}__offsets"] = off
        arrays[f"{name}__ngh"] = ngh
        arrays[f"{name}__src"] = np.array([src], np.uint32)
        arrays[f"{name}__levels"] = lv
        names.append(name)
    arrays["names"] = np.array(names)
    np.savez_compressed(os.path.join(HERE, "bfs_cases.npz"), **arrays)


def sha(path):
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


PAGE_GRAPHS = [
    # (name, generator args, pagesize MiB words)
    ("rmat_s16_ef16", {"kind": "rmat", "scale": 16, "ef": 16, "seed": 7}, 1),
    ("er_n200k_d4", {"kind": "er", "n": 200000, "m": 800000, "seed": 11}, 1),
]


def make_graph(spec):
    if spec["kind"] == "rmat":
        return _oracle.rmat_csr(spec["scale"], spec["ef"], seed=spec["seed"])
    return _oracle.er_csr(spec["n"], spec["m"], spec["seed"])


def write_page_hashes():
    tool = _oracle.REF_TOOL
    if not os.path.exists(tool):
        raise SystemExit("oracle/_ref/bin2pagecsr is not built (make -C oracle)")
    result = {}
    for name, spec, mb in PAGE_GRAPHS:
        off, ngh = make_graph(spec)
        n = len(off) - 1
        with tempfile.TemporaryDirectory() as td:
            base = os.path.join(td, "g")
            off.astype(np.uint32).tofile(base + "_csr.idx")  # the tool's default index_t is u32
            ngh.astype(np.uint32).tofile(base + "_csr.ngh")
            np.diff(off).astype(np.uint32).tofile(base + "_csr.deg")
            np.zeros(len(ngh), np.uint32).tofile(base + "_csr.wgh")
            subprocess.run([tool, base, str(mb), "u"], check=True, stdout=subprocess.DEVNULL)
            info = np.fromfile(base + "page.info", np.uint32).reshape(-1, 4)
            files = {"page.info": sha(base + "page.info"), "graph.info": sha(base + "graph.info")}
            for i in range(len(info)):
                files[f"page{i}"] = sha(base + f"page{i}")
            result[name] = {"spec": spec, "pagesize_mb": mb, "n": n, "m": int(len(ngh)),
                            "pages": info.tolist(), "sha256": files}
    with open(os.path.join(HERE, "pages.json"), "w") as f:
        json.dump(result, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    write_bfs_cases()
    write_page_hashes()
    print("golden fixtures written")
