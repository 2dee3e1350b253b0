"""CPU tests of the product's host side: C-ABI exports, formats, generators,
flags — each compared with the oracle or the reference tool's output."""
import ctypes
import hashlib
import json
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import scaph_amd
from scaph_amd import _lib
from tests import _oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
BIN = os.path.join(ROOT, "scaph_amd", "bin")


def declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(scaph_\w+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [("scaph_hip.h", "hip"), ("scaph_engine.h", "host")])
def test_c_abi_exports_every_declared_symbol(header, lib):
    L = _lib.hip() if lib == "hip" else _lib.host()
    names = declared_functions(header)
    assert len(names) > 10
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_host_page_cuts_match_oracle():
    off, ngh = O.rmat_csr(14, 16, seed=4)
    for P in (1 << 16, 1 << 18, 1 << 20):
        np.testing.assert_array_equal(scaph_amd.page_cuts(off, P), O.page_cuts(off, P))


def test_host_rmat_generator_equals_oracle():
    for scale, ef, seed in ((10, 8, 1), (13, 16, 9), (12, 4, 2)):
        a = scaph_amd.generate_rmat(scale, ef, seed=seed)
        b = O.rmat_csr(scale, ef, seed=seed)
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])
    a = scaph_amd.generate_rmat(11, 8, a=0.45, b=0.22, c=0.22, seed=5)
    b = O.rmat_csr(11, 8, a=0.45, b=0.22, c=0.22, seed=5)
    np.testing.assert_array_equal(a[1], b[1])


def test_host_er_generator_equals_oracle():
    a = scaph_amd.generate_er(3000, 24000, seed=17)
    b = O.er_csr(3000, 24000, 17)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


def sha(path):
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def test_host_page_files_match_reference_hashes():
    gold = json.load(open(os.path.join(GOLDEN, "pages.json")))
    for name, g in gold.items():
        spec = g["spec"]
        if spec["kind"] == "rmat":
            off, ngh = scaph_amd.generate_rmat(spec["scale"], spec["ef"], seed=spec["seed"])
        else:
            off, ngh = scaph_amd.generate_er(spec["n"], spec["m"], seed=spec["seed"])
        with tempfile.TemporaryDirectory() as td:
            base = os.path.join(td, "g")
            scaph_amd.write_graph(base, off, ngh, pagesize_words=g["pagesize_mb"] << 20,
                                  idx64=False, write_wgh=True)
            for fname, h in g["sha256"].items():
                assert sha(base + fname) == h, (name, fname)


def test_pagegencsr_cli_matches_reference_tool_bytes():
    """bin/pagegencsr (product) vs oracle/_ref/bin2pagecsr (reference source)."""
    if not os.path.exists(O.REF_TOOL):
        pytest.skip("reference tool not built here")
    off, ngh = O.rmat_csr(16, 32, seed=21)
    with tempfile.TemporaryDirectory() as td:
        for who in ("ref", "mine"):
            base = os.path.join(td, who)
            off.astype(np.uint32).tofile(base + "_csr.idx")
            ngh.tofile(base + "_csr.ngh")
            np.diff(off).astype(np.uint32).tofile(base + "_csr.deg")
            np.zeros(len(ngh), np.uint32).tofile(base + "_csr.wgh")
        subprocess.run([O.REF_TOOL, os.path.join(td, "ref"), "1", "u"], check=True,
                       stdout=subprocess.DEVNULL)
        subprocess.run([os.path.join(BIN, "pagegencsr"), os.path.join(td, "mine"), "1", "u"],
                       check=True, stdout=subprocess.DEVNULL)
        names = sorted(f[3:] for f in os.listdir(td) if f.startswith("ref") and "_csr" not in f)
        assert "page.info" in names and "graph.info" in names and "page1" in names
        for f in names:
            assert sha(os.path.join(td, "ref" + f)) == sha(os.path.join(td, "mine" + f)), f


def test_flag_syntax_gflags_forms():
    """-f=v, --f=v, -f v, -nobool, dash/underscore, hex, flagfile, unknown."""
    with tempfile.TemporaryDirectory() as td:
        ff = os.path.join(td, "flags")
        open(ff, "w").write("# comment\n-num_stream=7\n--threshold=4\n")
        args = ["-max-memory=2048", "--pagesize", "0x2", "-nocheck", "--check=yes",
                f"--flagfile={ff}", "-src", "5", "--", "-notaflag"]
        out = subprocess.run([os.path.join(BIN, "bfs"), *args, "-rounds=0"],
                             capture_output=True, text=True)
        # parsing succeeds; what fails next is the missing graph (or, on a
        # machine without a GPU, the device) -- never a flag
        assert "ERROR" not in out.stdout, out.stdout + out.stderr
        assert ("no -graphfile and no -gen" in out.stdout or "device" in out.stdout), out.stdout
        bad = subprocess.run([os.path.join(BIN, "bfs"), "-nosuchflag=1"], capture_output=True,
                             text=True)
        assert bad.returncode != 0 and "unknown command line flag 'nosuchflag'" in bad.stdout
        bad = subprocess.run([os.path.join(BIN, "bfs"), "-num_stream=abc"], capture_output=True,
                             text=True)
        assert bad.returncode != 0 and "illegal value" in bad.stdout


def test_engine_config_parse_via_capi_rejects_bad_threshold():
    # threshold > 16 would overflow the half-page compaction tail
    with pytest.raises(scaph_amd.ScaphError, match="threshold"):
        scaph_amd.BFS("-threshold=17")
    with pytest.raises(scaph_amd.ScaphError, match="page_direct"):
        scaph_amd.BFS("-page_direct=-1")


def test_host_cpu_bfs_equals_oracle():
    z = np.load(os.path.join(GOLDEN, "bfs_cases.npz"), allow_pickle=False)
    L = _lib.host()
    for name in z["names"]:
        off = z[f"{name}__offsets"]
        ngh = z[f"{name}__ngh"]
        src = int(z[f"{name}__src"][0])
        n = len(off) - 1
        out = np.empty(n, np.uint32)
        g = ngh if len(ngh) else np.zeros(1, np.uint32)
        assert L.scaph_host_cpu_bfs(n, off.ctypes.data, g.ctypes.data, src, 0xFFFFFFFF,
                                    out.ctypes.data) == 0
        np.testing.assert_array_equal(out, z[f"{name}__levels"])


# ------------------------------------------------------- page scheduler

def test_chunk_masks_match_oracle():
    L = _lib.host()
    Ob = O.lib()
    for f in ("scaph_host_max_zero_run",):
        getattr(L, f).argtypes = [ctypes.c_uint32]
        getattr(L, f).restype = ctypes.c_int
    L.scaph_host_fit_zero_run.argtypes = [ctypes.c_uint32, ctypes.c_int]
    L.scaph_host_fit_zero_run.restype = ctypes.c_int
    L.scaph_host_set_run.argtypes = [ctypes.c_uint32, ctypes.c_int, ctypes.c_int]
    L.scaph_host_set_run.restype = ctypes.c_uint32
    L.scaph_host_clear_run.argtypes = [ctypes.c_uint32, ctypes.c_int, ctypes.c_int]
    L.scaph_host_clear_run.restype = ctypes.c_uint32
    rng = np.random.default_rng(8)
    masks = list(range(1 << 14)) + [int(x) for x in rng.integers(0, 1 << 32, 30000, dtype=np.int64)]
    for m in masks:
        mi = int(np.uint32(m).view(np.int32))
        assert L.scaph_host_max_zero_run(m) == Ob.oracle_max_zero_run(mi)
        r = int(m % 32) + 1
        assert L.scaph_host_fit_zero_run(m, r) == Ob.oracle_fit_zero_run(mi, r), (hex(m), r)
        off, num = int(m % 29), int((m >> 5) % 4)
        assert L.scaph_host_set_run(m, off, num) == (Ob.oracle_set_run(mi, off, num) & 0xFFFFFFFF)
        assert L.scaph_host_clear_run(m, off, num) == (Ob.oracle_clear_run(mi, off, num) &
                                                       0xFFFFFFFF)


def _sim_lib():
    """tests/native/libpagehandle_sim.so: the test-only scheduler driver."""
    path = os.path.join(ROOT, "tests", "native", "libpagehandle_sim.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} missing: run __graft_entry__.build()")
    _lib.host()
    return ctypes.CDLL(path)


@pytest.mark.parametrize("predict", [0, 1])
@pytest.mark.parametrize("list_size,pool,thres,multi,seed", [
    (40, 8, 3, 1, 1), (40, 8, 3, 2, 2), (100, 10, 12, 1, 3), (16, 16, 4, 1, 4),
    (64, 3, 16, 3, 5), (12, 1, 8, 1, 6)])
def test_page_scheduler_invariants(list_size, pool, thres, multi, seed, predict):
    """Every active page runs at least once per iteration; a page runs only
    when resident; chunk runs in shared slots never overlap and match the
    slot masks (simulated executor, random completion order).  With a
    page-activity prediction (random values and keep-whole marks) the same
    holds, look-ahead copies never target a page of the running level, and
    every page over the threshold is still copied whole."""
    L = _sim_lib()
    L.scaph_host_pagehandle_simulate.argtypes = [ctypes.c_uint32, ctypes.c_uint32,
                                                 ctypes.c_uint64, ctypes.c_int, ctypes.c_uint32,
                                                 ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p,
                                                 ctypes.c_int]
    L.scaph_host_pagehandle_simulate.restype = ctypes.c_int
    rng = np.random.default_rng(seed)
    iters = 12
    # mix of inactive, cold (<= thres) and hot (> thres) pages, in chunk units
    datas = rng.choice([0, 0, 1, 2, thres, thres + 1, 32], size=(iters, list_size)).astype(np.uint32)
    stats = np.zeros(5, np.uint64)
    rc = L.scaph_host_pagehandle_simulate(list_size, pool, thres, multi, iters, datas.ctypes.data,
                                          seed, stats.ctypes.data, predict)
    assert rc == 0, f"invariant {rc} violated"
    assert stats[0] >= (datas > 0).sum()
    if predict and 3 <= pool < list_size:
        assert stats[4] > 0  # look-ahead copies were made


def test_committed_schedule_model_loads():
    from scaph_amd import learn as L
    p, mean, scale, gain = L.load_model()
    assert p.shape == (L.NPARAM,) and mean.shape == (16,) and scale.shape == (16,)
    w = L.bf16_to_f32(p)
    assert np.all(np.isfinite(w)) and np.all(scale > 0) and np.isfinite(gain)
    # bf16 round trip of the stored parameters is exact
    np.testing.assert_array_equal(L.f32_to_bf16(w), p)
    # the committed model reads the frontier's shares of the graph
    assert L.load_model_flags() & L.RELATIVE
    m = L.graph_mean(mean, L.RELATIVE, 1 << 20, 1 << 24)
    np.testing.assert_allclose(m[:3] - mean[:3], [20, 24, 24], rtol=1e-6)
    np.testing.assert_array_equal(m[3:], mean[3:])


def test_training_graph_specs():
    """The committed model's set: 200 seeded graphs, 70% R-MAT at scales
    18-24 and 30% Erdos-Renyi; --powerlaw trades a third of the ER share and
    a tenth of the R-MAT share for power-law digraphs.  The specs are a pure
    function of the seed."""
    from scaph_amd import learn as L
    base = L.training_specs(200)
    assert base == L.training_specs(200)
    kinds = [s["kind"] for s in base]
    assert kinds.count("rmat") == 140 and kinds.count("er") == 60
    assert all(18 <= s["scale"] <= 24 and 8 <= s["ef"] <= 32 and 0.45 <= s["a"] <= 0.65
               for s in base if s["kind"] == "rmat")
    assert all(8 * s["n"] <= s["m"] <= 64 * s["n"] for s in base if s["kind"] == "er")
    pl = [s["kind"] for s in L.training_specs(200, powerlaw=True)]
    assert (pl.count("rmat"), pl.count("er"), pl.count("zipf")) == (120, 40, 40)


def test_graphgen_weights_and_app_tools():
    """wmax= adds uniform weights on [1, wmax] without changing the edges;
    the wcc and sssp tools parse the shared flags."""
    with tempfile.TemporaryDirectory() as td:
        for tag, spec in (("w", "rmat:scale=11,ef=8,seed=3,wmax=50"),
                          ("u", "rmat:scale=11,ef=8,seed=3")):
            subprocess.run([os.path.join(BIN, "graphgen"), "-gen", spec, "-wgh", "-out",
                            os.path.join(td, tag)], check=True, stdout=subprocess.DEVNULL)
        for ext in ("_csr.idx", "_csr.ngh"):
            assert sha(os.path.join(td, "w" + ext)) == sha(os.path.join(td, "u" + ext))
        w = np.fromfile(os.path.join(td, "w_csr.wgh"), np.uint32)
        assert len(w) == len(np.fromfile(os.path.join(td, "w_csr.ngh"), np.uint32))
        assert w.min() == 1 and w.max() == 50
        assert abs(w.mean() - 25.5) < 1.0
        assert not np.fromfile(os.path.join(td, "u_csr.wgh"), np.uint32).any()
    for tool in ("wcc", "sssp"):
        out = subprocess.run([os.path.join(BIN, tool), "-src=3", "-check", "-rounds=2"],
                             capture_output=True, text=True)
        assert out.returncode != 0
        assert ("no -graphfile and no -gen" in out.stdout or "device" in out.stdout), out.stdout


def test_pagegencsr_transposed_pages():
    """pagegencsr ... t writes the in-edge graph (sources ascending per
    target) and its page files, cut like any other graph's."""
    off, ngh = O.rmat_csr(13, 16, seed=6)
    n = len(off) - 1
    src = np.repeat(np.arange(n), np.diff(off.astype(np.int64)))
    t_off, t_ngh = O.csr_from_edges(n, ngh, src)   # stable: sources stay ascending
    with tempfile.TemporaryDirectory() as td:
        base = os.path.join(td, "g")
        off.astype(np.uint64).tofile(base + "_csr.idx")
        ngh.tofile(base + "_csr.ngh")
        np.diff(off).astype(np.uint32).tofile(base + "_csr.deg")
        out = subprocess.run([os.path.join(BIN, "pagegencsr"), base, "1", "u", "t"],
                             capture_output=True, text=True, check=True)
        assert "in-edge pages:" in out.stdout
        np.testing.assert_array_equal(np.fromfile(base + "_in_csr.idx", np.uint64), t_off)
        np.testing.assert_array_equal(np.fromfile(base + "_in_csr.ngh", np.uint32), t_ngh)
        np.testing.assert_array_equal(np.fromfile(base + "_in_csr.deg", np.uint32),
                                      np.diff(t_off).astype(np.uint32))
        cuts = O.page_cuts(t_off, 1 << 20)
        info = np.fromfile(base + "_inpage.info", np.uint32).reshape(-1, 4)
        np.testing.assert_array_equal(info, cuts)
        assert len([f for f in os.listdir(td) if re.fullmatch(r"g_inpage\d+", f)]) == len(cuts)


def test_exporters_match_reference_tools():
    """bin2ligra[weight] / bin2totem[weight] (product) against the reference's
    exporters compiled from source (u32 offsets, the reference's only read
    width).  bin2totemweight of the reference overruns its weight array, so
    the weighted Totem file is checked as the unweighted reference file with
    the weights appended per line."""
    ref = os.path.join(ROOT, "oracle", "_ref")
    if not os.path.exists(os.path.join(ref, "bin2totem")):
        pytest.skip("reference exporters not built here")
    off, ngh = O.rmat_csr(12, 8, seed=14)
    rng = np.random.default_rng(3)
    wgh = rng.integers(0, 1 << 32, len(ngh), dtype=np.uint64).astype(np.uint32)
    with tempfile.TemporaryDirectory() as td:
        for who in ("ref", "mine"):
            base = os.path.join(td, who)
            off.astype(np.uint32).tofile(base + "_csr.idx")
            ngh.tofile(base + "_csr.ngh")
            np.diff(off).astype(np.uint32).tofile(base + "_csr.deg")
            wgh.tofile(base + "_csr.wgh")
        for tool in ("bin2ligra", "bin2ligraweight", "bin2totem"):
            subprocess.run([os.path.join(ref, tool), os.path.join(td, "ref")], check=True)
        for tool in ("bin2ligra", "bin2ligraweight", "bin2totem", "bin2totemweight"):
            subprocess.run([os.path.join(BIN, tool), os.path.join(td, "mine")], check=True)
        for suffix in (".ligra", ".ligra.weight", ".totem"):
            assert sha(os.path.join(td, "ref" + suffix)) == \
                sha(os.path.join(td, "mine" + suffix)), suffix
        ref_lines = open(os.path.join(td, "ref.totem")).read().splitlines()
        mine_lines = open(os.path.join(td, "mineweight.totem")).read().splitlines()
        assert mine_lines[:3] == ref_lines[:3]
        want = [f"{a}\t{w}" for a, w in zip(ref_lines[3:], wgh.tolist())]
        assert mine_lines[3:] == want
        # "w" as a second argument selects the weighted format as well
        os.remove(os.path.join(td, "mine.ligra.weight"))
        subprocess.run([os.path.join(BIN, "bin2ligra"), os.path.join(td, "mine"), "w"], check=True)
        assert sha(os.path.join(td, "ref.ligra.weight")) == \
            sha(os.path.join(td, "mine.ligra.weight"))


def test_pool_plan_equals_oracle_pool_size():
    """a15: the engine's page-pool planner (the -max_memory budget of
    src/app.cuh:281-286, 323-324) against the oracle's closed form, over
    budgets around every boundary: the vertex-data exit (-1), one slot, a
    partial pool and a pool that holds every page."""
    P = 10 << 20
    cases = []
    for n, pages in ((1 << 20, 2), (67_108_864, 220), (105_000_000, 389), (1000, 1), (0, 0),
                     (3_000_000_000, 7)):
        need = 9 * n + 24 * pages
        base = need >> 20
        for mib in (0, 1, base, base + 40, base + 41, base + 80, base + 81, base + 40 * pages,
                    base + 40 * pages + 41, 8192, 16142, 15258, 262144):
            cases.append((mib, n, pages, P))
    cases += [(16, 4096, 12, 4096), (1, 10, 3, 64), (2, 100000, 3, 1 << 16)]
    for mib, n, pages, P_ in cases:
        want = O.lib().oracle_pool_size(mib, n, pages, P_)
        got = scaph_amd.plan_pool(mib, n, pages, P_)
        assert got == want, (mib, n, pages, P_, got, want)
    # the configurations' budgets: C3's 16e9-byte pool and 8 GiB
    assert scaph_amd.plan_pool(16142, 105_000_000, 389, P) == 381
    assert scaph_amd.plan_pool(8192, 105_000_000, 389, P) == 182
    assert scaph_amd.plan_pool(100, 105_000_000, 389, P) == -1
