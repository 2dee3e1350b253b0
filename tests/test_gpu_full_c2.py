"""Full-size parity of the resident BFS on R-MAT scale 26, edge factor 32,
(a,b,c) = (.57,.19,.19), seed 2 (67.1 M vertices, 2.10 B edges): the
workload the benchmark times, with its own 64 giant-component sources.

  * fixed alpha/beta schedule: levels of the first four sources equal the
    oracle BFS bit for bit (the reference's check(), app/bfs.cu:67-117);
  * push-only, pull-only and fixed schedules on eight sources: the device
    invariant check finds no violated edge and no orphan, and the three
    schedules give the same levels;
  * the learned schedule (committed model) on the same graph: levels equal
    the oracle for the four sources.
"""
import os
import time

import numpy as np
import pytest

import bench
import scaph_amd
from scaph_amd import learn as L
from tests import _oracle as O

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]

SCALE, EF, SEED = 26, 32, 2


def say(*a):
    print("[gpu-test]", *a, flush=True)


@pytest.fixture(scope="module")
def c2():
    t0 = time.perf_counter()
    say("C2 build R-MAT", SCALE, EF, "seed", SEED)
    e = scaph_amd.BFS("-max_memory=262144", "-pagesize=10")
    e.gen_rmat_device(SCALE, EF, 0.57, 0.19, 0.19, seed=SEED, keep_host=True)
    info = e.info()
    assert info["n"] == 1 << SCALE and info["resident"] and info["pull"]
    srcs, tried = bench.pick_sources(e, info["n"], 64, seed=SEED, min_reach_frac=0.01)
    assert len(srcs) == 64
    off, ngh = e.host_csr()
    assert len(ngh) == info["m"]
    say("C2 graph", info, f"{time.perf_counter() - t0:.1f}s")
    want = {}
    for s in srcs[:4]:
        t1 = time.perf_counter()
        want[s] = O.bfs_mt(off, ngh, s)
        say("C2 oracle source", s, f"{time.perf_counter() - t1:.1f}s",
            "reached", int((want[s] != O.UNREACHED).sum()))
    yield {"e": e, "srcs": srcs, "want": want, "info": info}
    e.close()


def test_c2_fixed_schedule_equals_oracle(c2):
    e = c2["e"]
    e.set_schedule("fixed")
    for s, want in c2["want"].items():
        st = e.run(s)
        got = e.levels()
        bad = int((got != want).sum())
        say("C2 fixed source", s, "ms", round(st["ms"], 3), "depth", st["depth"], "mismatches", bad)
        assert bad == 0
        reached = want != O.UNREACHED
        assert st["n_reached"] == int(reached.sum())
        # the reference's capped form (MAXDEPTH 63, app/bfs.cu:18)
        assert np.array_equal(np.minimum(got, 63), np.minimum(want, 63))


def test_c2_schedules_agree_and_validate(c2):
    e = c2["e"]
    for s in c2["srcs"][:8]:
        ref = None
        for sched in ("fixed", "pull", "push"):
            e.set_schedule(sched)
            st = e.run(s)
            v = e.validate(s)
            say("C2 validate", sched, "source", s, "ms", round(st["ms"], 3), "violations", v)
            assert v == (0, 0)
            got = e.levels()
            if ref is None:
                ref = got
            else:
                assert np.array_equal(got, ref), f"{sched} differs from fixed for source {s}"
    e.set_schedule("fixed")


def test_c5_learned_schedule_equals_oracle(c2):
    """Config 5 on its held-out R-MAT graph: the committed model decides the
    directions, the levels stay the oracle's."""
    if not os.path.exists(L.DEFAULT_MODEL):
        pytest.fail("committed model scaph_amd/schedule_mlp.bin missing")
    e = c2["e"]
    L.model_for_graph(e, L.DEFAULT_MODEL)
    e.set_schedule("learned")
    try:
        for s, want in c2["want"].items():
            st = e.run(s)
            dirs = [r["dir"] for r in e.records()]
            bad = int((e.levels() != want).sum())
            say("C5 learned source", s, "ms", round(st["ms"], 3), "dirs", dirs, "mismatches", bad)
            assert bad == 0
    finally:
        e.set_schedule("fixed")
