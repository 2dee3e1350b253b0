"""Full-size parity of the batched BFS configurations.

  (a) multi-source bitset BFS on R-MAT scale 24, edge factor 16, seed 3:
      256 giant-component sources in one traversal, with the built-in and
      the autotuned launch choice — every query equals the single-source
      engine's levels, which the device invariant check accepts, and 16
      queries equal the oracle BFS bit for bit;
  (b) 4096 R-MAT graphs (scale 10..16, edge factor 8..32, seed 4), one
      source each, in one persistent launch, built-in and autotuned — every
      graph's levels equal the oracle BFS on that graph, and a sample of
      graphs equals the oracle's own R-MAT generator.
"""
import time

import numpy as np
import pytest

import bench
import scaph_amd
from tests import _oracle as O

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]


def say(*a):
    print("[gpu-test]", *a, flush=True)


def test_c4a_multi_source_256_queries_s24():
    t0 = time.perf_counter()
    with scaph_amd.BFS("-max_memory=262144", "-pagesize=10") as e:
        e.gen_rmat_device(24, 16, 0.57, 0.19, 0.19, seed=3, keep_host=True)
        info = e.info()
        srcs, _ = bench.pick_sources(e, info["n"], 256, seed=3, min_reach_frac=0.01)
        assert len(srcs) == 256
        src = np.array(srcs, np.uint32)
        say("C4a graph", info, f"{time.perf_counter() - t0:.1f}s")
        best, best_ms, log = e.msbfs_autotune(src)
        say("C4a autotuned", best, round(best_ms, 3), "candidates", len(log))
        off, ngh = e.host_csr()
        for tune in (None, best):
            st = e.msbfs(src, tune)
            assert st["queries"] == 256
            bad = []
            for q, s in enumerate(srcs):
                # the single-source run leaves the batch's state alone
                e.run(int(s))
                if tune is None:
                    assert e.validate(int(s)) == (0, 0), f"query {q}"
                single = e.levels()
                if tune is None and q % 16 == 0:
                    assert np.array_equal(single, O.bfs_mt(off, ngh, int(s))), f"query {q}"
                if not np.array_equal(e.msbfs_levels(q), single):
                    bad.append(q)
            say("C4a msbfs", tune, "ms", round(st["ms"], 3), "depth", st["depth"],
                "queries differing", len(bad))
            assert not bad


def test_c4b_4096_small_graphs_vs_oracle():
    t0 = time.perf_counter()
    with scaph_amd.Batch(count=4096, min_scale=10, max_scale=16, min_ef=8, max_ef=32,
                         seed=4) as b:
        info = b.info()
        say("C4b batch", info, f"{time.perf_counter() - t0:.1f}s")
        assert info["graphs"] == 4096
        best, best_ms, log = b.autotune()
        say("C4b autotuned", best, round(best_ms, 3), "candidates", len(log))
        scales = set()
        for tune in (None, best):
            ms = b.run(tune)
            t1 = time.perf_counter()
            bad = []
            for g in range(4096):
                G = b.graph(g)
                scales.add(G["scale"])
                off, ngh = b.csr(g)
                if not np.array_equal(b.levels(g), O.bfs(off, ngh, G["src"])):
                    bad.append(g)
            say("C4b run", tune, "ms", round(ms, 3), "graphs differing", len(bad),
                f"check {time.perf_counter() - t1:.1f}s")
            assert not bad, bad[:10]
        assert scales == set(range(10, 17))
        # the batch's graphs are the oracle generator's R-MAT graphs
        for g in range(0, 4096, 256):
            G = b.graph(g)
            off, ngh = b.csr(g)
            w_off, w_ngh = O.rmat_csr(G["scale"], G["edge_factor"], seed=G["seed"])
            assert np.array_equal(off, w_off) and np.array_equal(ngh, w_ngh), f"graph {g}"
