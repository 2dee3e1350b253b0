"""GPU parity tests of the engine's other App plugins — WCC (app/wcc.cu)
and SSSP (app/sssp.cu) — through the C-ABI, against the oracle bit for bit
on the same graphs: the committed BFS fixtures, R-MAT graphs (directed and
symmetric), out-lists long enough for the chunked pass, several pages, and
the error paths."""
import os

import numpy as np
import pytest

import scaph_amd
from tests import _oracle as O

pytestmark = pytest.mark.gpu

INF = O.SSSP_INF


def say(*a):
    print("[gpu-test]", *a, flush=True)


def engine(*flags):
    return scaph_amd.BFS("-max_memory=65536", *flags)


def fixtures():
    z = np.load(os.path.join(O.ROOT, "tests", "golden", "bfs_cases.npz"), allow_pickle=False)
    return [(str(n), z[f"{n}__offsets"], z[f"{n}__ngh"], int(z[f"{n}__src"][0]))
            for n in z["names"]]


def weighted_csr(n, src, dst, wgh):
    off, ngh = O.csr_from_edges(n, src, dst)
    order = np.argsort(np.asarray(src, np.int64), kind="stable")
    return off, ngh, np.asarray(wgh, np.uint32)[order]


def symmetric_rmat(scale, ef, seed):
    off, ngh = O.rmat_csr(scale, ef, seed=seed)
    n = len(off) - 1
    s = np.repeat(np.arange(n), np.diff(off.astype(np.int64)))
    return O.csr_from_edges(n, np.concatenate([s, ngh]), np.concatenate([ngh, s]))


@pytest.mark.parametrize("page_words", [0, 4096])
def test_wcc_fixtures(page_words):
    flags = [f"-page_words={page_words}"] if page_words else []
    for name, off, ngh, _ in fixtures():
        say("wcc fixture", name, "page_words", page_words)
        with engine(*flags) as e:
            e.load_csr(off, ngh)
            st = e.wcc()
            np.testing.assert_array_equal(e.levels(), O.wcc(off, ngh), err_msg=name)
            assert st["depth"] >= 1


@pytest.mark.parametrize("scale,ef,seed,sym", [(12, 8, 3, False), (14, 16, 5, True),
                                               (16, 16, 1, False)])
def test_wcc_rmat(scale, ef, seed, sym):
    off, ngh = symmetric_rmat(scale, ef, seed) if sym else O.rmat_csr(scale, ef, seed=seed)
    say("wcc rmat", scale, ef, seed, "symmetric" if sym else "directed")
    want = O.wcc(off, ngh)
    with engine() as e:
        e.load_csr(off, ngh)
        st = e.wcc()
        np.testing.assert_array_equal(e.levels(), want)
        # every round relaxes the out-lists of its active vertices; the first
        # round has every vertex active
        assert st["m_reached"] >= len(ngh)


def test_wcc_long_lists_chunked():
    """Out-lists longer than one wave item (4096 edges) continue as chunks:
    a hub with 20000 out-edges whose label must reach all of them, behind a
    chain that lowers the hub's label only in later rounds."""
    k = 20000
    n = k + 10
    hub = n - 1
    src = [hub] * k + [0, 1, 2, 3]
    dst = list(range(5, 5 + k)) + [1, 2, 3, hub]
    off, ngh = O.csr_from_edges(n, src, dst)
    want = O.wcc(off, ngh)
    assert want[5] == 0 and want[5 + k - 1] == 0
    with engine() as e:
        e.load_csr(off, ngh)
        e.wcc()
        np.testing.assert_array_equal(e.levels(), want)


@pytest.mark.parametrize("page_words", [0, 4096])
def test_sssp_fixtures_unit_and_random_weights(page_words):
    flags = [f"-page_words={page_words}"] if page_words else []
    rng = np.random.default_rng(21)
    for name, off, ngh, src in fixtures():
        say("sssp fixture", name, "page_words", page_words)
        for wgh in (np.ones(len(ngh), np.uint32), rng.integers(0, 50, len(ngh)).astype(np.uint32)):
            with engine(*flags) as e:
                e.load_csr(off, ngh)
                e.set_weights(wgh)
                e.sssp(src)
                np.testing.assert_array_equal(e.levels(), O.sssp(off, ngh, wgh, src),
                                              err_msg=name)


@pytest.mark.parametrize("scale,ef,seed,wmax,page_words", [(12, 8, 2, 1000, 0),
                                                           (14, 16, 4, 64, 65536),
                                                           (16, 16, 6, 255, 0)])
def test_sssp_rmat(scale, ef, seed, wmax, page_words):
    off, ngh = O.rmat_csr(scale, ef, seed=seed)
    rng = np.random.default_rng(seed)
    wgh = rng.integers(1, wmax + 1, len(ngh)).astype(np.uint32)
    deg = np.diff(off.astype(np.int64))
    flags = [f"-page_words={page_words}"] if page_words else []
    with engine(*flags) as e:
        e.load_csr(off, ngh)
        e.set_weights(wgh)
        for src in (int(np.argmax(deg)), int(np.nonzero(deg)[0][3])):
            say("sssp rmat", scale, ef, seed, "src", src)
            e.sssp(src)
            np.testing.assert_array_equal(e.levels(), O.sssp(off, ngh, wgh, src))


def test_sssp_long_lists_and_chain():
    """A 20000-edge out-list (chunked) whose cheap edges improve distances
    already set through a long chain."""
    k = 20000
    chain = 300
    n = k + chain + 2
    hub = n - 1
    src, dst, w = [], [], []
    for i in range(chain):        # 0 -> 1 -> ... -> chain, weight 1
        src.append(i), dst.append(i + 1), w.append(1)
    src.append(0), dst.append(hub), w.append(7)
    for j in range(k):
        src.append(hub), dst.append(chain + 1 + j % (k - 1)), w.append(1 + j % 13)
    for i in range(chain + 1, chain + 200):
        src.append(i), dst.append(i - chain), w.append(0)  # shortcut back into the chain
    off, ngh, wgh = weighted_csr(n, src, dst, w)
    want = O.sssp(off, ngh, wgh, 0)
    with engine() as e:
        e.load_csr(off, ngh)
        e.set_weights(wgh)
        st = e.sssp(0)
        np.testing.assert_array_equal(e.levels(), want)
        assert st["depth"] >= 2


def test_apps_then_bfs_and_errors():
    off, ngh = O.rmat_csr(12, 8, seed=9)
    deg = np.diff(off.astype(np.int64))
    s = int(np.argmax(deg))
    with engine() as e:
        e.load_csr(off, ngh)
        with pytest.raises(RuntimeError):
            e.sssp(s)                       # no weights loaded
        with pytest.raises(RuntimeError):
            e.set_weights(np.ones(len(ngh) + 1, np.uint32))   # wrong count
        e.set_weights(np.full(len(ngh), 2, np.uint32))
        with pytest.raises(RuntimeError):
            e.sssp(len(off))                # source out of range
        e.wcc()
        e.sssp(s)
        lv = O.bfs(off, ngh, s)
        np.testing.assert_array_equal(e.levels(), np.where(lv == O.UNREACHED, INF, 2 * lv))
        e.run(s)                            # BFS on the same engine afterwards
        np.testing.assert_array_equal(e.levels(), lv)
        e.wcc()
        np.testing.assert_array_equal(e.levels(), O.wcc(off, ngh))


@pytest.mark.parametrize("flags", [["-num_stream=1"], ["-num_stream=3"],
                                   ["-num_stream=3", "-multi=2"], ["-num_stream=5", "-multi=3"],
                                   ["-num_stream=2", "-threshold=16"], ["-threshold=0"],
                                   ["-threshold=16", "-multi=2", "-num_stream=3"]])
def test_apps_paged(flags):
    """WCC and SSSP out of core: a pool of a few page slots, pages copied
    whole or compacted to their active Nodes (with their weights), on one or
    several exec streams, and a BFS run in between (which replaces slots
    without weights)."""
    off, ngh = O.rmat_csr(16, 16, seed=8)
    rng = np.random.default_rng(4)
    wgh = rng.integers(1, 200, len(ngh)).astype(np.uint32)
    deg = np.diff(off.astype(np.int64))
    base = ["-page_words=262144", "-max_memory=3", "-threshold=3"]
    with scaph_amd.BFS(*base, *flags) as e:
        e.load_csr(off, ngh)
        assert not e.info()["resident"] and e.info()["pages"] >= 4
        say("paged apps", flags)
        st = e.wcc()
        np.testing.assert_array_equal(e.levels(), O.wcc(off, ngh))
        assert st["bytes_h2d"] > 0
        e.set_weights(wgh)
        srcs = (int(np.argmax(deg)), int(np.nonzero(deg)[0][7]))
        e.sssp(srcs[0])
        np.testing.assert_array_equal(e.levels(), O.sssp(off, ngh, wgh, srcs[0]))
        e.run(srcs[1])
        np.testing.assert_array_equal(e.levels(), O.bfs(off, ngh, srcs[1]))
        st2 = e.sssp(srcs[1])
        np.testing.assert_array_equal(e.levels(), O.sssp(off, ngh, wgh, srcs[1]))
        if "-threshold=0" in flags:
            assert st["pages_compacted"] == 0 and st2["pages_compacted"] == 0
        elif "-threshold=16" in flags:
            assert st["pages_compacted"] + st2["pages_compacted"] > 0


@pytest.mark.parametrize("tool,extra", [("wcc", []), ("sssp", ["-src=7"]),
                                        ("sssp", ["-src=7", "-page_words=8192"]),
                                        ("sssp", ["-src=7", "-page_words=65536",
                                                  "-max_memory=1"])])
def test_app_tools_check_passes(tool, extra):
    """bin/wcc and bin/sssp on a generated weighted graph, with the app's own
    CPU check (label sweeps / Dijkstra) and two rounds."""
    import subprocess
    exe = os.path.join(O.ROOT, "scaph_amd", "bin", tool)
    say("tool", tool, extra)
    out = subprocess.run([exe, "-gen=rmat:scale=14,ef=16,seed=2,wmax=100", "-check", "-rounds=2",
                          *extra], capture_output=True, text=True, timeout=100)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "check passed!" in out.stdout and f"{tool} cost" in out.stdout


def test_apps_edge_cases():
    """Single vertex, no edges, self-loops only, parallel edges with different
    weights, zero weights, an isolated source, and weights near 2^24."""
    cases = []
    cases.append(("single", 1, [], [], []))
    cases.append(("no edges", 50, [], [], []))
    cases.append(("self loops", 20, list(range(20)), list(range(20)), [5] * 20))
    cases.append(("parallel", 4, [0, 0, 0, 1, 1, 2], [1, 1, 1, 2, 2, 3], [9, 3, 7, 4, 1, 0]))
    rng = np.random.default_rng(31)
    s = rng.integers(0, 300, 2000)
    d = rng.integers(0, 300, 2000)
    cases.append(("big weights", 300, s, d, rng.integers(1 << 23, 1 << 24, 2000)))
    cases.append(("zero weights", 300, s, d, np.zeros(2000, np.int64)))
    for name, n, src, dst, w in cases:
        say("edge case", name)
        off, ngh, wgh = weighted_csr(n, src, dst, w)
        with engine() as e:
            e.load_csr(off, ngh)
            e.wcc()
            np.testing.assert_array_equal(e.levels(), O.wcc(off, ngh), err_msg=name)
            e.set_weights(wgh)
            for source in sorted({0, n - 1}):
                e.sssp(source)
                np.testing.assert_array_equal(e.levels(), O.sssp(off, ngh, wgh, source),
                                              err_msg=f"{name} {source}")


@pytest.mark.parametrize("delta", [1, 37, 300, 1 << 20])
def test_sssp_near_far_rounds(delta):
    """-sssp_delta: only active vertices below a moving bound relax each
    round; the distances equal the oracle's for any bucket width (fixtures,
    zero weights, R-MAT, a long chain behind a hub)."""
    rng = np.random.default_rng(delta)
    graphs = []
    for name, off, ngh, src in fixtures():
        graphs.append((name, off, ngh, rng.integers(0, 60, len(ngh)).astype(np.uint32), src))
    off, ngh = O.rmat_csr(15, 16, seed=3)
    deg = np.diff(off.astype(np.int64))
    graphs.append(("rmat15", off, ngh, rng.integers(1, 256, len(ngh)).astype(np.uint32),
                   int(np.argmax(deg))))
    n = 3000
    s = list(range(n - 1)) + [0] * 500
    d = list(range(1, n)) + list(rng.integers(1, n, 500))
    w = [1] * (n - 1) + list(rng.integers(1000, 5000, 500))
    off, ngh, wgh = weighted_csr(n, s, d, w)
    graphs.append(("chain", off, ngh, wgh, 0))
    for name, off, ngh, wgh, src in graphs:
        say("near/far", delta, name)
        with engine(f"-sssp_delta={delta}") as e:
            e.load_csr(off, ngh)
            e.set_weights(wgh)
            st = e.sssp(src)
            np.testing.assert_array_equal(e.levels(), O.sssp(off, ngh, wgh, src),
                                          err_msg=f"{name} delta {delta}")
            assert st["depth"] >= 1
