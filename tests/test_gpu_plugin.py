"""A third-party app on the App<> plugin surface (the reference's
src/app.cuh:72-135 with app/bfs.cu:39-65): tests/native/plugin_app.cpp
subclasses relax_app, initialises the engine's device values and active
bitmap in its gpu_init through the kernel C-ABI, and lets the engine run
min-relaxing rounds.  Its result — for every vertex the least index of a
seed that reaches it — is checked against the oracle BFS from each seed,
resident and out of core."""
import ctypes
import os

import numpy as np
import pytest

from scaph_amd import _lib
from tests import _oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
U = 0xFFFFFFFF


def plugin():
    path = os.path.join(ROOT, "tests", "native", "libplugin_app.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} missing: run __graft_entry__.build()")
    _lib.host()
    L = ctypes.CDLL(path)
    P = ctypes.c_void_p
    L.plugin_seed_labels.argtypes = [ctypes.c_uint32, P, P, P, ctypes.c_uint32, ctypes.c_char_p,
                                     P, ctypes.POINTER(ctypes.c_uint64), ctypes.c_char_p,
                                     ctypes.c_size_t]
    L.plugin_seed_labels.restype = ctypes.c_int
    return L


def seed_labels(off, ngh, seeds, flags):
    n = len(off) - 1
    off = np.ascontiguousarray(off, np.uint64)
    ngh = np.ascontiguousarray(ngh, np.uint32)
    seeds = np.ascontiguousarray(seeds, np.uint32)
    out = np.empty(n, np.uint32)
    rounds = ctypes.c_uint64()
    err = ctypes.create_string_buffer(1024)
    rc = plugin().plugin_seed_labels(n, off.ctypes.data, ngh.ctypes.data, seeds.ctypes.data,
                                     len(seeds), flags.encode(), out.ctypes.data,
                                     ctypes.byref(rounds), err, 1024)
    _lib.check(rc, err.value.decode())
    return out, rounds.value


def oracle_labels(off, ngh, seeds):
    want = np.full(len(off) - 1, U, np.uint32)
    for i, s in enumerate(seeds):
        reach = O.bfs(off, ngh, int(s)) != U
        want[reach] = np.minimum(want[reach], i)
    return want


@pytest.mark.parametrize("flags", ["-max_memory=65536",
                                   "-max_memory=2 -page_words=65536 -build_pull=false",
                                   "-max_memory=3 -page_words=65536 -threshold=12 -num_stream=2"])
def test_seed_label_plugin_app(flags):
    off, ngh = O.rmat_csr(16, 8, seed=41)
    rng = np.random.default_rng(7)
    n = len(off) - 1
    seeds = rng.integers(0, n, 12).astype(np.uint32)
    seeds[3] = seeds[9]  # a repeated seed keeps its lower index
    got, rounds = seed_labels(off, ngh, seeds, flags)
    print("[gpu-test] plugin", flags, "rounds", rounds, flush=True)
    np.testing.assert_array_equal(got, oracle_labels(off, ngh, seeds))
    assert rounds >= 1
