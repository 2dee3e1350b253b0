"""GPU parity tests: levels from the HIP path (through the C-ABI) against
the oracle, bit for bit, for every schedule; plus the device graph builder
and the per-page count kernel."""
import ctypes
import os

import numpy as np
import pytest

import scaph_amd
from scaph_amd import _lib
from tests import _oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
U = 0xFFFFFFFF
SCHEDULES = ["fixed", "push", "pull"]


def say(*a):
    print("[gpu-test]", *a, flush=True)


def engine(*flags):
    if any(f.startswith("-max_memory") for f in flags):
        return scaph_amd.BFS(*flags)
    return scaph_amd.BFS("-max_memory=65536", *flags)


def cases():
    z = np.load(os.path.join(GOLDEN, "bfs_cases.npz"), allow_pickle=False)
    return [(str(n), z[f"{n}__offsets"], z[f"{n}__ngh"], int(z[f"{n}__src"][0]),
             z[f"{n}__levels"]) for n in z["names"]]


@pytest.mark.parametrize("page_words", [0, 4096])
def test_golden_fixtures_all_schedules(page_words):
    flags = [f"-page_words={page_words}"] if page_words else []
    for name, off, ngh, src, want in cases():
        say("fixture", name, "page_words", page_words)
        with engine(*flags) as e:
            e.load_csr(off, ngh)
            for sched in SCHEDULES:
                e.set_schedule(sched)
                st = e.run(src)
                got = e.levels()
                np.testing.assert_array_equal(got, want, err_msg=f"{name} {sched}")
                reached = want != U
                assert st["n_reached"] == reached.sum()
                deg = np.diff(off.astype(np.int64))
                assert st["m_reached"] == deg[reached].sum()


@pytest.mark.parametrize("scale,ef,seed,page_words", [(12, 8, 3, 0), (16, 16, 1, 0),
                                                      (14, 32, 5, 65536)])
def test_rmat_parity_with_oracle(scale, ef, seed, page_words):
    off, ngh = O.rmat_csr(scale, ef, seed=seed)
    deg = np.diff(off.astype(np.int64))
    srcs = [int(np.argmax(deg)), int(np.nonzero(deg)[0][7]), 0]
    flags = [f"-page_words={page_words}"] if page_words else []
    with engine(*flags) as e:
        e.load_csr(off, ngh)
        for src in srcs:
            want = O.bfs(off, ngh, src)
            for sched in SCHEDULES:
                say("rmat", scale, ef, seed, "src", src, sched)
                e.set_schedule(sched)
                e.run(src)
                np.testing.assert_array_equal(e.levels(), want, err_msg=f"src {src} {sched}")
            assert e.check(src) == 0


@pytest.mark.parametrize("head", [["-pull_head=true", "-pull_head_outdeg=false"],
                                  ["-pull_head=false"],
                                  ["-pull_head=true", "-pull_head_outdeg=true"],
                                  ["-pull_head_live=false"],
                                  ["-pull_head_live=false", "-pull_head_outdeg=false",
                                   "-pull_prefetch=false"],
                                  ["-pull_prefetch=false"],
                                  ["-pull_head_index=false"],
                                  ["-pull_head_index=true", "-pull_head_live=false"],
                                  ["-pull_head_index=true", "-pull_prefetch=false"],
                                  ["-pull_head_index=true", "-pull_head_outdeg=false"]])
def test_pull_head_records(head):
    """The pull expand with and without the 16-byte in-list head records,
    on the host-loaded and the device-built graph; indexed records (the
    list's page offset beside its first two entries; in-degrees of 255 and
    more read the Node) with and without the live-only numbering."""
    off, ngh = O.rmat_csr(15, 16, seed=8)
    deg = np.diff(off.astype(np.int64))
    srcs = [int(np.argmax(deg)), int(np.nonzero(deg)[0][3])]
    with engine(*head) as e:
        e.load_csr(off, ngh)
        for src in srcs:
            e.set_schedule("pull")
            e.run(src)
            np.testing.assert_array_equal(e.levels(), O.bfs(off, ngh, src))
    with engine(*head) as e:
        e.gen_rmat_device(15, 16, seed=8)
        for src in srcs:
            for sched in ("pull", "fixed"):
                e.set_schedule(sched)
                e.run(src)
                np.testing.assert_array_equal(e.levels(), O.bfs(off, ngh, src))


def test_indexed_head_records_builder():
    """scaph_list_heads_indexed against a numpy restatement: {e0, e1,
    (offset of the list in its page's edges) << 8 | min(length, 255),
    out-degree}, for every vertex and for the vertices with a nonempty list
    only (numbered per 1024-id window), on pages of 4096 words with lists of
    0, 1, 2 and more than 255 entries."""
    hip = _lib.hip()
    off, ngh = O.rmat_csr(13, 16, seed=3)
    n = len(off) - 1
    deg = np.diff(off.astype(np.int64))
    assert (deg == 0).any() and (deg == 1).any() and (deg >= 255).any()
    pages = O.page_cuts(off, 4096)
    assert len(pages) > 8
    outdeg = np.random.default_rng(2).integers(0, 1 << 20, n).astype(np.uint32)
    page_of = np.searchsorted(pages[:, 0].astype(np.int64), np.arange(n), side="right") - 1
    idx = off[:-1].astype(np.int64) - off[pages[page_of, 0]].astype(np.int64)
    ng = ngh.astype(np.int64)
    o = off[:-1].astype(np.int64)
    want = np.empty((n, 4), np.uint32)
    want[:, 0] = np.where(deg > 0, ng[np.minimum(o, len(ng) - 1)], 0xFFFFFFFF)
    want[:, 1] = np.where(deg > 1, ng[np.minimum(o + 1, len(ng) - 1)], 0xFFFFFFFF)
    want[:, 2] = (idx << 8) | np.minimum(deg, 255)
    want[:, 3] = outdeg
    words = ((n + 63) // 64) * 2
    bits = np.ones(words * 32, np.uint8)
    bits[:n] = deg == 0
    no_in = np.packbits(bits, bitorder="little").view(np.uint32)
    live = deg > 0
    windows = (words + 31) // 32
    base = np.array([live[:1024 * k].sum() for k in range(windows)], np.uint32)
    bufs = []

    def dev(arr):
        p = ctypes.c_void_p()
        _lib.check(hip.scaph_dev_alloc(ctypes.byref(p), max(arr.nbytes, 8)))
        _lib.check(hip.scaph_copy_h2d_async(p, arr.ctypes.data, arr.nbytes, None))
        bufs.append(p)
        return p

    try:
        d_off, d_ngh = dev(np.ascontiguousarray(off, np.uint64)), dev(ngh)
        d_out, d_pages = dev(outdeg), dev(np.ascontiguousarray(pages))
        d_no, d_base = dev(no_in), dev(base)
        for live_only in (False, True):
            rows = int(live.sum()) if live_only else n
            got = np.zeros((rows, 4), np.uint32)
            d_head = dev(got)
            _lib.check(hip.scaph_list_heads_indexed(d_off, d_ngh, n, d_out, d_pages, len(pages),
                                                    d_no if live_only else None,
                                                    d_base if live_only else None, d_head, None))
            _lib.check(hip.scaph_device_sync())
            _lib.check(hip.scaph_copy_d2h_async(got.ctypes.data, d_head, got.nbytes, None))
            _lib.check(hip.scaph_device_sync())
            np.testing.assert_array_equal(got, want[live] if live_only else want,
                                          err_msg=f"live_only={live_only}")
    finally:
        for p in bufs:
            hip.scaph_dev_free(p)


def test_rmat_s20_ef16_seed1_from_max_degree():
    """Config 1: R-MAT s20 ef16 (a,b,c)=(.57,.19,.19) seed 1, source = max out-degree."""
    off, ngh = O.rmat_csr(20, 16, seed=1)
    deg = np.diff(off.astype(np.int64))
    src = int(np.argmax(deg))
    want = O.bfs(off, ngh, src)
    want63 = O.bfs(off, ngh, src, inf=63)
    with engine() as e:
        e.load_csr(off, ngh)
        for sched in SCHEDULES:
            say("rmat s20 ef16 seed 1", sched)
            e.set_schedule(sched)
            st = e.run(src)
            got = e.levels()
            np.testing.assert_array_equal(got, want)
            np.testing.assert_array_equal(np.where(got <= 62, got, 63), want63)
            assert e.validate(src) == (0, 0)
            say("rmat s20 ef16 seed 1", sched, "ms", round(st["ms"], 3), "depth", st["depth"])


def test_many_pages_global_page_lookup():
    """More pages than the LDS page table holds (global lookup path)."""
    off, ngh = O.er_csr(200000, 1600000, 21)
    src = 17
    want = O.bfs(off, ngh, src)
    with engine("-page_words=1024") as e:
        e.load_csr(off, ngh)
        assert e.info()["pages"] > 1024
        for sched in SCHEDULES:
            say("many pages", sched)
            e.set_schedule(sched)
            e.run(src)
            np.testing.assert_array_equal(e.levels(), want)


def test_device_builder_equals_oracle_generator():
    for scale, ef, seed in ((12, 16, 2), (15, 8, 9)):
        say("device builder", scale, ef, seed)
        off, ngh = O.rmat_csr(scale, ef, seed=seed)
        with engine() as e:
            e.gen_rmat_device(scale, ef, seed=seed, keep_host=True)
            hoff, hngh = e.host_csr()
            np.testing.assert_array_equal(hoff, off)
            np.testing.assert_array_equal(hngh, ngh)
            np.testing.assert_array_equal(e.out_degrees(), np.diff(off).astype(np.uint32))
            src = int(np.argmax(np.diff(off.astype(np.int64))))
            e.run(src)
            np.testing.assert_array_equal(e.levels(), O.bfs(off, ngh, src))


def test_edge_cases():
    say("edge cases")
    with engine() as e:
        # one vertex, no edges
        e.load_csr(np.array([0, 0], np.uint64), np.zeros(0, np.uint32))
        e.run(0)
        np.testing.assert_array_equal(e.levels(), [0])
        # no edges at all, source in the middle
        e.load_csr(np.zeros(11, np.uint64), np.zeros(0, np.uint32))
        e.run(4)
        want = np.full(10, U, np.uint32)
        want[4] = 0
        np.testing.assert_array_equal(e.levels(), want)
        # source with out-degree 0 in a non-trivial graph
        off, ngh = O.csr_from_edges(5, [0, 1, 2], [1, 2, 3])
        e.load_csr(off, ngh)
        e.run(3)
        np.testing.assert_array_equal(e.levels(), [U, U, U, 0, U])
        # out-of-range source is an error, not a fault
        with pytest.raises(scaph_amd.ScaphError):
            e.run(5)


def test_page_counts_kernel_vs_oracle():
    """scaph_page_counts (get_page_info/get_total_info) on a random frontier."""
    hip = _lib.hip()
    off, ngh = O.er_csr(40000, 320000, 12)
    n = 40000
    deg = np.diff(off).astype(np.uint32)
    pages = O.page_cuts(off, 1 << 14)
    rng = np.random.default_rng(4)
    stat = (rng.random(n) < 0.15).astype(np.uint8)
    words = ((n + 63) // 64) * 2
    bits = np.zeros(words * 32, np.uint8)
    bits[:n] = stat
    bitmap = np.packbits(bits, bitorder="little").view(np.uint32)
    want_nodes, want_datas, want_tot = O.page_counts(pages, stat, deg)
    bufs = {}

    def dev(arr):
        p = ctypes.c_void_p()
        _lib.check(hip.scaph_dev_alloc(ctypes.byref(p), max(arr.nbytes, 8)))
        _lib.check(hip.scaph_copy_h2d_async(p, arr.ctypes.data, arr.nbytes, None))
        bufs[id(arr)] = p
        return p

    d_pages, d_front, d_deg = dev(np.ascontiguousarray(pages)), dev(bitmap), dev(deg)
    out_n = np.zeros(len(pages), np.uint32)
    out_d = np.zeros(len(pages), np.uint32)
    tot = np.zeros(2, np.uint64)
    d_n, d_d, d_t = dev(out_n), dev(out_d), dev(tot)
    for weighted in (0, 1):
        _lib.check(hip.scaph_page_counts(d_pages, len(pages), d_front, d_deg, weighted, d_n, d_d,
                                         d_t, None))
        _lib.check(hip.scaph_device_sync())
        for arr, p in ((out_n, d_n), (out_d, d_d), (tot, d_t)):
            _lib.check(hip.scaph_copy_d2h_async(arr.ctypes.data, p, arr.nbytes, None))
        _lib.check(hip.scaph_device_sync())
        wn, wd, wt = O.page_counts(pages, stat, deg, weighted=bool(weighted))
        np.testing.assert_array_equal(out_n, wn)
        np.testing.assert_array_equal(out_d, wd)
        np.testing.assert_array_equal(tot, wt)
    for p in bufs.values():
        hip.scaph_dev_free(p)


def test_frontier_bitmap_to_queue_ragged():
    """scaph_frontier_bitmap_to_queue on vertex counts around its block span
    (8192 words), its 32-word thread runs and the word edge (bits past n in
    the last word are set and must be ignored): the list holds exactly the
    frontier vertices, each once, in increasing order within a block."""
    hip = _lib.hip()
    rng = np.random.default_rng(21)
    span = 8192 * 32
    for n, dens in ((1, 1.0), (31, 0.5), (33, 0.5), (65536, 0.001), (65537 + 31, 0.3),
                    (span * 3 + 5, 0.02), (span * 2 + 32 * 5 + 7, 0.4), (1000003, 0.6),
                    (4194321, 0.0)):
        say("bitmap_to_queue", n, dens)
        words = ((n + 63) // 64) * 2
        bits = np.zeros(words * 32, np.uint8)
        bits[:n] = rng.random(n) < dens
        bits[n:] = 1  # garbage past n
        bitmap = np.packbits(bits, bitorder="little").view(np.uint32)
        want = np.nonzero(bits[:n])[0].astype(np.uint32)
        d_bm, d_q, d_c = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        _lib.check(hip.scaph_dev_alloc(ctypes.byref(d_bm), bitmap.nbytes))
        _lib.check(hip.scaph_dev_alloc(ctypes.byref(d_q), max(4 * n, 8)))
        cnt = _lib.LevelCounters()
        _lib.check(hip.scaph_dev_alloc(ctypes.byref(d_c), ctypes.sizeof(cnt)))
        try:
            _lib.check(hip.scaph_copy_h2d_async(d_bm, bitmap.ctypes.data, bitmap.nbytes, None))
            _lib.check(hip.scaph_copy_h2d_async(d_c, ctypes.addressof(cnt), ctypes.sizeof(cnt),
                                                None))
            _lib.check(hip.scaph_frontier_bitmap_to_queue(d_bm, n, d_q, d_c, None))
            _lib.check(hip.scaph_device_sync())
            _lib.check(hip.scaph_copy_d2h_async(ctypes.addressof(cnt), d_c, ctypes.sizeof(cnt),
                                                None))
            got = np.zeros(len(want), np.uint32)
            if len(want):
                _lib.check(hip.scaph_copy_d2h_async(got.ctypes.data, d_q, got.nbytes, None))
            _lib.check(hip.scaph_device_sync())
        finally:
            for p_ in (d_bm, d_q, d_c):
                hip.scaph_dev_free(p_)
        assert cnt.q_len == len(want)
        np.testing.assert_array_equal(np.sort(got), want)
        # each block's slice is in increasing order
        blk = got // span
        for b in np.unique(blk):
            seg = got[blk == b]
            assert np.all(np.diff(seg.astype(np.int64)) > 0)


@pytest.mark.parametrize("page_words,max_mib,threshold,multi,streams", [
    (65536, 1, 3, 1, 5), (65536, 1, 16, 1, 2), (16384, 1, 0, 1, 5), (16384, 1, 8, 2, 3),
    (8192, 1, 12, 1, 1)])
def test_paged_out_of_core_parity(page_words, max_mib, threshold, multi, streams):
    """Out-of-core mode: device pool smaller than the page set, pages stream
    from pinned memory (full copies and compacted cold pages; no direct
    pushes from host memory here)."""
    off, ngh = O.rmat_csr(14, 16, seed=6)
    deg = np.diff(off.astype(np.int64))
    with engine(f"-page_words={page_words}", f"-max_memory={max_mib}",
                f"-threshold={threshold}", f"-multi={multi}", f"-num_stream={streams}",
                "-page_direct=0") as e:
        e.load_csr(off, ngh)
        inf = e.info()
        assert not inf["resident"], inf
        for src in (int(np.argmax(deg)), int(np.nonzero(deg)[0][3])):
            say("paged", page_words, max_mib, threshold, multi, streams, "src", src)
            want = O.bfs(off, ngh, src)
            st = e.run(src)
            np.testing.assert_array_equal(e.levels(), want)
            assert st["page_execs"] >= 1 and st["bytes_h2d"] > 0 and st["pages_direct"] == 0
            if threshold == 0:
                assert st["pages_compacted"] == 0
            assert e.check(src) == 0


@pytest.mark.parametrize("flags", [[], ["-page_groups=0"], ["-page_groups=0", "-page_hint=false"],
                                   ["-push_small=1"], ["-push_small=1000000"],
                                   ["-push_small=1", "-push_rows=false"], ["-page_groups=1"],
                                   ["-page_groups=64", "-multi=2"],
                                   ["-push_small=1", "-push_scan=1"],
                                   ["-push_small=1", "-push_xcd=1"],
                                   ["-page_direct=4096"], ["-page_direct=1099511627776"],
                                   ["-page_direct=1099511627776", "-push_small=1"]])
def test_paged_page_push_variants(flags):
    """Page pushes sized by the page counts: grouped launches over several
    resident pages, one-launch small pages and groups, full pages read
    through their frontier bits (pages of ~14K Nodes, so two 8192-Node
    segments, at unaligned first vertices), compacted pages, and the
    per-page and unhinted paths, and uncached pages pushed straight from
    pinned host memory (-page_direct: some, every one, through the one-launch
    small path too), all against the oracle."""
    off, ngh = O.rmat_csr(16, 16, seed=8)
    deg = np.diff(off.astype(np.int64))
    direct = any(f.startswith("-page_direct") for f in flags)
    base = [] if direct else ["-page_direct=0"]
    with engine("-page_words=262144", "-max_memory=3", "-threshold=3", "-num_stream=3",
                *base, *flags) as e:
        e.load_csr(off, ngh)
        inf = e.info()
        assert not inf["resident"] and inf["pages"] >= 4, inf
        for src in (int(np.argmax(deg)), int(np.nonzero(deg)[0][5])):
            say("paged variants", flags, "src", src)
            want = O.bfs(off, ngh, src)
            st = e.run(src)
            np.testing.assert_array_equal(e.levels(), want, err_msg=f"{flags} {src}")
            say("  ", {k: st[k] for k in ("bytes_h2d", "pages_direct", "pages_compacted")})
            if direct:
                assert st["pages_direct"] > 0
                if "1099511627776" in flags[0]:
                    assert st["pages_compacted"] == 0
            else:
                assert st["bytes_h2d"] > 0 and st["pages_direct"] == 0


def test_pull_heavy_in_lists():
    """In-lists far longer than one pull work chunk, with the only frontier
    in-neighbour at the very end (and a hub with no frontier in-neighbour)."""
    k = 20000
    n = k + 4
    s_, a, t, h = 0, n - 1, 1, 2      # source, its out-neighbour (largest id), targets
    src = [s_, a]
    dst = [a, t]
    xs = np.arange(3, 3 + k)
    src += list(xs) + list(xs)
    dst += [t] * k + [h] * k           # t: k + 1 in-edges (a last), h: k in-edges, none reachable
    off, ngh = O.csr_from_edges(n, src, dst)
    want = O.bfs(off, ngh, s_)
    assert want[t] == 2 and want[h] == O.UNREACHED
    with engine() as e:
        e.load_csr(off, ngh)
        for sched in SCHEDULES:
            say("heavy in-lists", sched)
            e.set_schedule(sched)
            e.run(s_)
            np.testing.assert_array_equal(e.levels(), want)


@pytest.mark.parametrize("flags", [["-push_rows=false"], ["-push_rows=true", "-push_grid=64"],
                                   ["-pull_grid=16"], ["-push_small=1"], ["-push_small=1000000"],
                                   ["-pull_prefetch=false"], ["-pull_waves=4"],
                                   ["-pull_waves=4", "-pull_prefetch=false", "-pull_grid=7"],
                                   ["-pull_simd=8"], ["-pull_simd=8", "-pull_prefetch=false"],
                                   ["-pull_simd=8", "-pull_waves=4"], ["-pull_hubs=0"],
                                   ["-pull_hubs=1000"], ["-pull_hubs=229376", "-pull_waves=4"],
                                   ["-pull_hubs=262144"], ["-pull_sparse=0"], ["-pull_sparse=1"],
                                   ["-pull_phases=1"], ["-pull_phases=2", "-pull_prefetch=false"],
                                   ["-pull_phases=2", "-pull_head_live=false", "-pull_waves=4"],
                                   ["-pull_phases=2", "-pull_head_index=false", "-pull_hubs=0"],
                                   ["-pull_phases=2", "-pull_head=false"],
                                   ["-pull_phases=2", "-pull_head_outdeg=false"],
                                   ["-pull_phases=3"], ["-pull_phases=3", "-pull_prefetch=false"],
                                   ["-pull_phases=3", "-pull_head_live=false", "-pull_waves=4"],
                                   ["-pull_phases=3", "-pull_head_index=false", "-pull_hubs=0"],
                                   ["-pull_phases=3", "-pull_head_outdeg=false"],
                                   ["-pull_phases=3", "-pull_simd=8", "-pull_hubs=1000"],
                                   ["-pull_sparse=1", "-pull_head_live=false"],
                                   ["-pull_sparse=1", "-pull_head=false"],
                                   ["-pull_sparse=1", "-pull_head_index=false", "-pull_hubs=0"],
                                   ["-pull_summary=6", "-pull_hubs=1000"],
                                   ["-pull_summary=5", "-pull_summary_div=0", "-pull_hubs=1000"],
                                   ["-pull_summary=9", "-pull_summary_div=0", "-pull_phases=1",
                                    "-pull_hubs=0"],
                                   ["-push_coop=1"], ["-push_coop=1000000000"],
                                   ["-push_coop=1", "-push_small=1"],
                                   ["-push_coop=300", "-push_small=1"],
                                   ["-push_small=1", "-push_scan=1"],
                                   ["-push_scan=1000000000"], ["-push_small=1", "-push_xcd=1"],
                                   ["-push_small=1", "-push_xcd=1", "-push_rows=false",
                                    "-push_grid=24"]])
def test_launch_shape_variants(flags):
    """The push/pull launch-shape switches change timing only."""
    off, ngh = O.rmat_csr(16, 16, seed=4)
    deg = np.diff(off.astype(np.int64))
    srcs = [int(np.argmax(deg)), int(np.nonzero(deg)[0][9])]
    with engine(*flags) as e:
        e.load_csr(off, ngh)
        for src in srcs:
            want = O.bfs(off, ngh, src)
            for sched in SCHEDULES:
                e.set_schedule(sched)
                e.run(src)
                np.testing.assert_array_equal(e.levels(), want, err_msg=f"{flags} {src} {sched}")


@pytest.mark.parametrize("flags", [["-pull_summary=5"], ["-pull_summary=6"],
                                   ["-pull_summary=7", "-pull_head=false"],
                                   ["-pull_summary=10", "-pull_hubs=0"],
                                   ["-pull_summary=6", "-pull_head_live=false", "-pull_waves=4"],
                                   ["-pull_summary=5", "-pull_phases=3"],
                                   ["-pull_summary=8", "-pull_phases=3", "-pull_head_outdeg=false"]])
def test_pull_frontier_summary(flags):
    """The pull with the frontier summary (one bit per 2^shift vertices)
    tested before each non-hub probe, forced on every pull level, on vertex
    counts that are not a multiple of the summary's span: levels equal the
    oracle's.  Few hubs (graphs this small would otherwise make every source
    a hub, and the summary guards non-hub probes only)."""
    rng = np.random.default_rng(11)
    graphs = [O.rmat_csr(15, 16, seed=5)]
    for n in (100003, 70001):
        m = 12 * n
        src = rng.integers(0, n, m)
        # a skewed in-side: low ids collect many in-edges, so the in-lists
        # mix hubs and long non-hub runs
        dst = (rng.pareto(1.2, m) * 50).astype(np.int64) % n
        graphs.append(O.csr_from_edges(n, src, dst))
    hubs = [] if any(f.startswith("-pull_hubs") for f in flags) else ["-pull_hubs=500"]
    with engine(*flags, *hubs, "-pull_summary_div=0") as e:
        for off, ngh in graphs:
            e.load_csr(off, ngh)
            deg = np.diff(off.astype(np.int64))
            for src in (int(np.argmax(deg)), int(np.nonzero(deg)[0][5])):
                want = O.bfs(off, ngh, src)
                for sched in ("pull", "fixed"):
                    say("summary", flags, len(off) - 1, src, sched)
                    e.set_schedule(sched)
                    e.run(src)
                    np.testing.assert_array_equal(e.levels(), want,
                                                  err_msg=f"{flags} {src} {sched}")


def test_xcd_round_robin_dispatch():
    """The sliced push expand gives target eighth x to blocks b with b mod 8
    == x, which keeps that eighth in one XCD's L2 when workgroups go to the
    eight XCDs round-robin (block b on XCD (s + b) mod 8, the start s
    varying between launches).  This checks that dispatch order on the
    device (it affects speed only; every tile has one owner either way)."""
    hip = _lib.hip()
    blocks = 4096
    out = np.full(blocks, 99, np.uint32)
    p = ctypes.c_void_p()
    _lib.check(hip.scaph_dev_alloc(ctypes.byref(p), out.nbytes))
    try:
        _lib.check(hip.scaph_xcc_ids(p, blocks, None))
        _lib.check(hip.scaph_copy_d2h_async(out.ctypes.data, p, out.nbytes, None))
        _lib.check(hip.scaph_device_sync())
    finally:
        hip.scaph_dev_free(p)
    say("xcc of blocks 0..15", out[:16].tolist())
    assert out.max() < 16
    xcds = int(out.max()) + 1
    if xcds == 8:
        want = (int(out[0]) + np.arange(blocks, dtype=np.uint32)) % 8
        np.testing.assert_array_equal(out, want)


def hw_id_tuples(hip, stream, blocks=65536):
    out = np.zeros(2 * blocks, np.uint32)
    p = ctypes.c_void_p()
    _lib.check(hip.scaph_dev_alloc(ctypes.byref(p), out.nbytes))
    try:
        _lib.check(hip.scaph_hw_ids(p, blocks, stream))
        _lib.check(hip.scaph_stream_sync(stream))
        _lib.check(hip.scaph_copy_d2h_async(out.ctypes.data, p, out.nbytes, None))
        _lib.check(hip.scaph_device_sync())
    finally:
        hip.scaph_dev_free(p)
    xcc, hw = out[0::2], out[1::2]
    # (XCC, shader engine, shader array, CU) of each workgroup
    return set(zip(xcc.tolist(), ((hw >> 8) & 0xFF).tolist()))


def test_cu_limited_stream_reaches_fewer_cus():
    """-tricknum's CU mask: launches on a stream that withholds CUs reach at
    most the CUs left (the reference's trick kernel occupies that many SMs
    with a spinning block instead)."""
    hip = _lib.hip()
    props = _lib.DeviceProps()
    _lib.check(hip.scaph_get_device_props(0, ctypes.byref(props)))
    cus = props.compute_units
    counts = {}
    for reserved in (0, cus // 2, cus - 8):
        s = ctypes.c_void_p()
        _lib.check(hip.scaph_stream_create_cu_limited(ctypes.byref(s), reserved))
        try:
            counts[reserved] = len(hw_id_tuples(hip, s))
        finally:
            hip.scaph_stream_destroy(s)
    say("CUs", cus, "distinct CUs reached by reserved count", counts)
    assert counts[0] <= cus
    assert counts[cus // 2] <= cus - cus // 2
    assert counts[cus - 8] <= 8
    assert counts[cus - 8] < counts[cus // 2] < counts[0]
    s = ctypes.c_void_p()
    assert hip.scaph_stream_create_cu_limited(ctypes.byref(s), cus) != 0


@pytest.mark.parametrize("flags", [["-tricknum=200"],
                                   ["-tricknum=128", "-page_words=262144", "-max_memory=3",
                                    "-threshold=3", "-num_stream=3"]])
def test_tricknum_bfs_parity(flags):
    off, ngh = O.rmat_csr(16, 16, seed=8)
    deg = np.diff(off.astype(np.int64))
    src = int(np.argmax(deg))
    say("tricknum", flags)
    with engine(*flags) as e:
        e.load_csr(off, ngh)
        e.run(src)
        np.testing.assert_array_equal(e.levels(), O.bfs(off, ngh, src))


def test_graphfile_with_transposed_pages():
    """-graphfile with the transposed-page generator's files: the engine takes
    the stored in-edge graph (after checking it is the transpose) and the
    levels stay exact; a stored transpose that differs is refused."""
    import subprocess
    import tempfile
    off, ngh = O.rmat_csr(14, 16, seed=12)
    deg = np.diff(off.astype(np.int64))
    src = int(np.argmax(deg))
    want = O.bfs(off, ngh, src)
    pagegen = os.path.join(ROOT, "scaph_amd", "bin", "pagegencsr")
    with tempfile.TemporaryDirectory() as td:
        base = os.path.join(td, "g")
        scaph_amd.write_graph(base, off, ngh, pagesize_words=0, idx64=True, write_wgh=False)
        subprocess.run([pagegen, base, "1", "u", "t"], check=True, stdout=subprocess.DEVNULL)
        assert os.path.exists(base + "_in_csr.ngh")
        for sched in ("fixed", "pull"):
            say("graphfile with transposed pages", sched)
            with engine(f"-graphfile={base}", f"-schedule={sched}") as e:
                e.load()
                e.run(src)
                np.testing.assert_array_equal(e.levels(), want)
        t_ngh = np.fromfile(base + "_in_csr.ngh", np.uint32)
        t_ngh[len(t_ngh) // 2] ^= 1
        t_ngh.tofile(base + "_in_csr.ngh")
        with engine(f"-graphfile={base}") as e:
            with pytest.raises(RuntimeError, match="not the transpose"):
                e.load()


def test_memory_budget_pool_equals_oracle():
    """a15 on the device: the pool the engine allocates for -max_memory is
    the oracle's pool size (src/app.cuh:281-286, 323-324), the graph is
    resident exactly when the pool holds every page, levels are right at
    every pool size, and a budget the vertex arrays exhaust fails with the
    reference's message (src/app.cuh:310-313)."""
    off, ngh = O.rmat_csr(16, 16, seed=6)
    n = len(off) - 1
    P = 65536
    pages = len(O.page_cuts(off, P))
    assert pages > 8
    deg = np.diff(off.astype(np.int64))
    src = int(np.argmax(deg))
    want = O.bfs(off, ngh, src)
    for mib in (1, 2, 3, 5, 64):
        expect = O.lib().oracle_pool_size(mib, n, pages, P)
        with engine(f"-max_memory={mib}", f"-page_words={P}", "-build_pull=false") as e:
            e.load_csr(off, ngh)
            got = e.pool()
            say("budget", mib, "MiB: pool", got, "oracle", expect, "pages", pages)
            assert got["slots"] == expect
            assert e.info()["resident"] == (expect == pages)
            e.run(src)
            np.testing.assert_array_equal(e.levels(), want)
    with engine("-max_memory=0", f"-page_words={P}", "-build_pull=false") as e:
        with pytest.raises(scaph_amd.ScaphError, match="vertex data out of gpu memory"):
            e.load_csr(off, ngh)
