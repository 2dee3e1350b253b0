"""GPU parity of the batched BFS paths: the multi-source bitset BFS (every
query's levels against the oracle BFS from its source) and the batch of
small graphs in one persistent launch."""
import numpy as np
import pytest

import scaph_amd
from tests import _oracle as O

pytestmark = pytest.mark.gpu

PUSH_ALWAYS = {"push_div": 1, "push_mdiv": 1}
PULL_ALWAYS = {"push_div": 1 << 30, "push_mdiv": 1 << 30}


def say(*a):
    print("[gpu-test]", *a, flush=True)


def _sources(off, count, seed):
    rng = np.random.default_rng(seed)
    n = len(off) - 1
    deg = np.diff(off.astype(np.int64))
    src = rng.integers(0, n, count).astype(np.uint32)
    # a few repeats and the highest-degree vertex
    if count > 4:
        src[1] = src[0]
        src[2] = int(np.argmax(deg))
    return src


def _check_all(e, off, ngh, src, label):
    for q, s in enumerate(src):
        want = O.bfs(off, ngh, int(s))
        got = e.msbfs_levels(q)
        if not np.array_equal(got, want):
            bad = np.nonzero(got != want)[0][:5]
            raise AssertionError(f"{label}: query {q} (source {s}) differs at {bad}: "
                                 f"got {got[bad]}, want {want[bad]}")


@pytest.mark.parametrize("tune", [
    None, PUSH_ALWAYS, PULL_ALWAYS,
    {"pull_block": 1024, "use_graph": 1, "segment": 3},
    {"pull_block": 512, "use_graph": 1, "segment": 8, "push_div": 16, "push_mdiv": 8},
    {"pull_block": 256, "pull_grid": 8},
    {"persistent": 1}, {"persistent": 1, **PUSH_ALWAYS}, {"persistent": 1, **PULL_ALWAYS}])
def test_msbfs_256_queries_vs_oracle(tune):
    off, ngh = O.rmat_csr(14, 16, seed=21)
    src = _sources(off, 256, 5)
    with scaph_amd.BFS("-max_memory=65536") as e:
        e.load_csr(off, ngh)
        say("msbfs 256 queries", tune)
        st = e.msbfs(src, tune)
        assert st["queries"] == 256 and st["depth"] >= 2
        _check_all(e, off, ngh, src, str(tune))
        # pairs reached = sum over queries of reached vertices
        reached = sum(int((O.bfs(off, ngh, int(s)) != O.UNREACHED).sum()) for s in src[:8])
        assert st["pairs"] >= reached
        cnt = e.msbfs_counters()
        assert cnt[0]["bits"] == 256 and cnt[st["depth"]]["n_active"] == 0


@pytest.mark.parametrize("flags", [["-pull_hubs=0"], ["-pull_hubs=300"], ["-pull_hubs=229376"]])
def test_msbfs_hub_counts(flags):
    """Pull levels with no hubs, a few hubs and every vertex a hub (the
    hubs' frontier words gathered per level): every query equals the
    oracle."""
    off, ngh = O.rmat_csr(14, 16, seed=24)
    src = _sources(off, 256, 6)
    with scaph_amd.BFS("-max_memory=65536", *flags) as e:
        e.load_csr(off, ngh)
        say("msbfs hub counts", flags)
        for tune in (None, PULL_ALWAYS):
            e.msbfs(src, tune)
            _check_all(e, off, ngh, src, f"{flags} {tune}")


@pytest.mark.parametrize("count", [1, 37, 64, 65, 200])
def test_msbfs_partial_batches(count):
    off, ngh = O.rmat_csr(13, 8, seed=22)
    src = _sources(off, count, count)
    with scaph_amd.BFS("-max_memory=65536") as e:
        e.load_csr(off, ngh)
        say("msbfs partial", count)
        for tune in (None, {"use_graph": 1, "segment": 4}):
            e.msbfs(src, tune)
            _check_all(e, off, ngh, src, f"count={count} {tune}")


def test_msbfs_many_pages_and_heavy_out_lists():
    """More pages than the LDS page table holds (Erdos-Renyi, 64-word pages),
    out-lists longer than one push chunk (always-push) and long in-lists on
    the whole wave (always-pull) on R-MAT."""
    cases = [("er", O.er_csr(60000, 480000, 24), "-page_words=64"),
             ("rmat", O.rmat_csr(15, 32, seed=23), "-page_words=0")]
    for name, (off, ngh), flags in cases:
        deg = np.diff(off.astype(np.int64))
        if name == "rmat":
            assert deg.max() > 2048
        src = _sources(off, 64, 6)
        with scaph_amd.BFS("-max_memory=65536", flags) as e:
            e.load_csr(off, ngh)
            if name == "er":
                assert e.info()["pages"] > 1024
            for tune in (PUSH_ALWAYS, PULL_ALWAYS, None, {"persistent": 1, **PUSH_ALWAYS},
                         {"persistent": 1, **PULL_ALWAYS}):
                say("msbfs", name, flags, tune)
                e.msbfs(src, tune)
                _check_all(e, off, ngh, src, f"{name} {tune}")


def test_msbfs_edge_cases():
    # isolated vertices, a path, a self loop, duplicate sources
    edges = [(0, 1), (1, 2), (2, 3), (3, 3), (5, 4), (4, 5)]
    n = 8
    off, ngh = O.csr_from_edges(n, [a for a, _ in edges], [b for _, b in edges])
    src = np.array([0, 0, 3, 7, 4, 6], np.uint32)
    with scaph_amd.BFS("-max_memory=65536") as e:
        e.load_csr(off, ngh)
        for tune in (None, PUSH_ALWAYS, PULL_ALWAYS, {"use_graph": 1, "segment": 2},
                     {"persistent": 1}, {"persistent": 1, **PULL_ALWAYS}):
            st = e.msbfs(src, tune)
            _check_all(e, off, ngh, src, f"edge cases {tune}")
        with pytest.raises(scaph_amd.ScaphError):
            e.msbfs(np.array([n], np.uint32))
        with pytest.raises(scaph_amd.ScaphError):
            e.msbfs(np.zeros(257, np.uint32))


def test_msbfs_deeper_than_level_buffers():
    """A path longer than the 32 level buffers: every mode reports it (the
    persistent launch stops at the last buffer), and a shorter path after it
    runs normally."""
    n = 40
    off, ngh = O.csr_from_edges(n, list(range(n - 1)), list(range(1, n)))
    with scaph_amd.BFS("-max_memory=65536") as e:
        e.load_csr(off, ngh)
        for tune in (None, {"use_graph": 1, "segment": 4}, {"persistent": 1}):
            say("msbfs deep path", tune)
            with pytest.raises(scaph_amd.ScaphError):
                e.msbfs(np.array([0], np.uint32), tune)
            e.msbfs(np.array([20, 35], np.uint32), tune)
            _check_all(e, off, ngh, np.array([20, 35], np.uint32), f"short path {tune}")


def _check_batch_graph(b, g):
    G = b.graph(g)
    off, ngh = b.csr(g)
    # the batch's graph g is the R-MAT graph of its seed, scale and edge factor
    w_off, w_ngh = O.rmat_csr(G["scale"], G["edge_factor"], seed=G["seed"])
    np.testing.assert_array_equal(off, w_off)
    np.testing.assert_array_equal(ngh, w_ngh)
    np.testing.assert_array_equal(b.levels(g), O.bfs(off, ngh, G["src"]), err_msg=f"graph {g}")


@pytest.mark.parametrize("tune", [None, {"block": 1024, "wgs_per_cu": 1},
                                  {"block": 512, "wgs_per_cu": 8, "use_graph": 1},
                                  {"block": 256, "wgs_per_cu": 2, "use_graph": 1}])
def test_small_graph_batch_vs_oracle(tune):
    with scaph_amd.Batch(count=96, min_scale=4, max_scale=12, min_ef=1, max_ef=24, seed=7) as b:
        info = b.info()
        assert info["graphs"] == 96
        say("batch", info, tune)
        b.run(tune)
        for g in range(96):
            _check_batch_graph(b, g)
        # a second run (graph replay or not) gives the same levels
        before = [b.levels(g) for g in range(0, 96, 7)]
        b.run(tune)
        for i, g in enumerate(range(0, 96, 7)):
            np.testing.assert_array_equal(b.levels(g), before[i])


def test_small_graph_batch_largest_scale_and_autotune():
    with scaph_amd.Batch(count=24, min_scale=15, max_scale=16, min_ef=8, max_ef=32, seed=9) as b:
        best, ms, log = b.autotune()
        say("batch autotune", best, ms, len(log))
        assert len(log) == 24 and ms > 0
        b.run()  # the tuned choice
        for g in range(0, 24, 5):
            _check_batch_graph(b, g)


def test_batch_rejects_bad_specs():
    with pytest.raises(scaph_amd.ScaphError):
        scaph_amd.Batch(count=4, min_scale=10, max_scale=17)
    with pytest.raises(scaph_amd.ScaphError):
        scaph_amd.Batch(count=0)
