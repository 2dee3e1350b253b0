"""scaph_bfs_push_noalloc (the push level whose workspace is a fixed device
array, for callers that cannot allocate) against scaph_bfs_push with a
caller workspace, through the C-ABI: two calls in a row, the second with a
smaller frontier, each checked bit for bit against the other entry point
and against the oracle's next level; and the refusal of a frontier whose
workspace exceeds the array."""
import ctypes

import numpy as np
import pytest

import scaph_amd
from scaph_amd import _lib
from tests import _oracle as O

pytestmark = pytest.mark.gpu

U = 0xFFFFFFFF
SCRATCH_TOO_SMALL = 0x5C500001


def say(*a):
    print("[gpu-test]", *a, flush=True)


class Dev:
    """Device buffers owned by the test (freed at exit)."""

    def __init__(self):
        self.hip = _lib.hip()
        self.bufs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        _lib.check(self.hip.scaph_dev_alloc(ctypes.byref(p), max(int(nbytes), 4)))
        self.bufs.append(p)
        return p

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        _lib.check(self.hip.scaph_copy_h2d_async(p, arr.ctypes.data, arr.nbytes, None))
        return p

    def get(self, p, dtype, count):
        out = np.empty(count, dtype)
        _lib.check(self.hip.scaph_copy_d2h_async(out.ctypes.data, p, out.nbytes, None))
        _lib.check(self.hip.scaph_device_sync())
        return out

    def free(self):
        _lib.check(self.hip.scaph_device_sync())
        for p in self.bufs:
            self.hip.scaph_dev_free(p)
        self.bufs = []


def state_entering(lv, L, n):
    """Level array, visited bitmap and frontier queue entering level L."""
    level = np.where(lv <= L, lv, U).astype(np.uint32)
    words = ((n + 63) // 64) * 2
    bits = np.zeros(words * 32, np.uint8)
    bits[:n] = lv <= L
    visited = np.packbits(bits, bitorder="little").view(np.uint32)
    queue = np.nonzero(lv == L)[0].astype(np.uint32)
    return level, visited, queue


def test_push_noalloc_matches_push_with_workspace():
    off, ngh = O.rmat_csr(16, 16, seed=4)
    n = len(off) - 1
    deg = np.diff(off.astype(np.int64))
    src = int(np.argmax(deg))
    lv = O.bfs(off, ngh, src)
    hip = _lib.hip()
    cap = hip.scaph_bfs_push_noalloc_max_bytes()
    assert cap == 64 << 20
    with scaph_amd.BFS("-max_memory=65536", "-build_pull=false") as e:
        e.load_csr(off, ngh)
        desc = e.graph_desc(False)
        degree = ctypes.c_void_p(e.device_degree())
        d = Dev()
        try:
            ws_bytes = hip.scaph_push_workspace_bytes(n, int(len(ngh)))
            ws = d.alloc(ws_bytes)
            sizes = []
            for L in (2, 1):  # the second call has the smaller frontier
                level, visited, queue = state_entering(lv, L, n)
                n_in = len(queue)
                m_in = int(deg[queue].sum())
                sizes.append(n_in)
                assert m_in > 8192  # past the one-launch path: the workspace is used
                assert hip.scaph_push_workspace_bytes(n_in, m_in) <= cap
                q_in = d.put(queue)
                out = {}
                for which in ("workspace", "noalloc"):
                    dl, dv = d.put(level), d.put(visited)
                    q_out = d.alloc(4 * n)
                    cnt = d.put(np.zeros(6, np.uint64))
                    say("push level", L, "n_in", n_in, "m_in", m_in, which)
                    if which == "workspace":
                        rc = hip.scaph_bfs_push(ctypes.byref(desc), degree, dl, dv, q_in, n_in,
                                                m_in, q_out, cnt, L, ws, ws_bytes, None, None)
                    else:
                        rc = hip.scaph_bfs_push_noalloc(ctypes.byref(desc), degree, dl, dv, q_in,
                                                        n_in, m_in, q_out, cnt, L, None, None)
                    _lib.check(rc)
                    _lib.check(hip.scaph_device_sync())
                    c = d.get(cnt, np.uint64, 6)
                    out[which] = (d.get(dl, np.uint32, n), d.get(dv, np.uint32, len(visited)),
                                  np.sort(d.get(q_out, np.uint32, n)[:int(c[2])]), c)
                a, b = out["workspace"], out["noalloc"]
                np.testing.assert_array_equal(a[0], b[0])
                np.testing.assert_array_equal(a[1], b[1])
                np.testing.assert_array_equal(a[2], b[2])
                np.testing.assert_array_equal(a[3][:4], b[3][:4])  # n_next m_next q_len scanned
                # and both are the oracle's next level
                want_level, want_visited, _ = state_entering(lv, L + 1, n)
                np.testing.assert_array_equal(b[0], want_level)
                np.testing.assert_array_equal(b[1][:len(want_visited)], want_visited)
                np.testing.assert_array_equal(b[2], np.nonzero(lv == L + 1)[0].astype(np.uint32))
                assert int(b[3][0]) == int((lv == L + 1).sum())
            assert sizes[1] < sizes[0]
            # a frontier whose workspace exceeds the array is refused up front
            rc = hip.scaph_bfs_push_noalloc(ctypes.byref(desc), degree, ws, ws, ws, n,
                                            1 << 40, ws, ws, 0, None, None)
            assert rc == SCRATCH_TOO_SMALL
            assert b"no-alloc" in hip.scaph_error_string(rc)
        finally:
            d.free()
