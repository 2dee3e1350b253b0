"""The power-law generator on the device against the oracle's, and the
out-of-core BFS on such a graph (pages streamed through the page pool)
against the oracle BFS."""
import numpy as np
import pytest

import scaph_amd
from tests import _oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


def test_device_zipf_graph_equals_oracle():
    n, x_min, alpha, cap, seed = 120000, 1.2, 1.9, 20000, 3
    w_off, w_ngh = O.zipf_csr(n, x_min, alpha, cap, seed)
    with scaph_amd.BFS("-max_memory=65536", "-build_pull=false") as e:
        say("gen_zipf", n)
        e.gen_zipf(n, x_min, alpha, cap, seed)
        off, ngh = e.host_csr()
        np.testing.assert_array_equal(off, w_off)
        np.testing.assert_array_equal(ngh, w_ngh)


@pytest.mark.parametrize("max_mib,threshold,direct", [(2, 3, 0), (3, 12, 0), (2, 3, 65536),
                                                     (3, 12, 1 << 40)])
def test_out_of_core_bfs_on_power_law_graph(max_mib, threshold, direct):
    """Paged BFS on a power-law digraph: cold pages copied (whole or
    compacted) with -page_direct=0, or pushed from pinned host memory."""
    n, cap = 120000, 20000
    off, ngh = O.zipf_csr(n, 1.2, 1.9, cap, 4)
    deg = np.diff(off.astype(np.int64))
    with scaph_amd.BFS(f"-max_memory={max_mib}", "-page_words=65536", f"-threshold={threshold}",
                       "-build_pull=false", f"-page_direct={direct}") as e:
        e.gen_zipf(n, 1.2, 1.9, cap, 4)
        info = e.info()
        say("paged zipf", info)
        assert not info["resident"]
        for src in (int(np.argmax(deg)), int(np.nonzero(deg)[0][7])):
            st = e.run(src)
            np.testing.assert_array_equal(e.levels(), O.bfs(off, ngh, src))
            assert st["page_execs"] >= st["page_execs_cached"]
            if direct == 0:
                assert st["bytes_h2d"] > 0 and st["pages_direct"] == 0
            else:
                assert st["pages_direct"] > 0
