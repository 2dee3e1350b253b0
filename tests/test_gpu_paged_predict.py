"""The paged scheduler's page-activity prediction (-page_predict): levels
stay the oracle's whatever is predicted, and on a graph whose page reuse
order matters the learned prediction (and the exact-knowledge bound) copy
fewer bytes over PCIe than the plain first-free-slot policy.

The graph (pages of 65536 words, a pool of 9 slots for 14 pages):
  page 0      the source (out-edges to every A vertex) and filler vertices
              no edge reaches;
  pages 1-5   B: reached at level 3 (from C);
  pages 6-8   A: reached at level 1 (from the source);
  pages 9-13  C: reached at level 2 (from A).
The pool starts with pages 0-8.  At level 2 the five C pages need slots:
the plain policy takes free slots in slot order (page 0, then B pages, which
level 3 needs again); the prediction knows pages 0 and 6-8 can never be
active again (no unvisited vertex with in-edges) and reuses those first,
then copies the one evicted B page ahead while level 2 runs.
"""
import numpy as np
import pytest

import scaph_amd
from scaph_amd import learn as L
from tests import _oracle as O

pytestmark = pytest.mark.gpu

P = 65536
PER_PAGE = 4095  # vertices of out-degree 13 per page: 16 * 4095 < 65536
DEG = 13


def say(*a):
    print("[gpu-test]", *a, flush=True)


def layered_graph():
    nA, nB, nC = 3 * PER_PAGE, 5 * PER_PAGE, 5 * PER_PAGE
    fill = (P - 1 - nA) // 3 - 1  # page 0: source + fillers, 3(1 + fill) + nA < P
    s0 = 0
    b0 = 1 + fill
    a0 = b0 + nB
    c0 = a0 + nA
    n = c0 + nC
    src, dst = [np.zeros(nA, np.int64)], [a0 + np.arange(nA)]
    j = np.arange(DEG)
    bi = np.arange(nB)[:, None]
    src.append(np.repeat(b0 + np.arange(nB), DEG))
    dst.append((a0 + (bi * DEG + j) % nA).ravel())
    ai = np.arange(nA)[:, None]
    src.append(np.repeat(a0 + np.arange(nA), DEG))
    dst.append((c0 + (ai * DEG + j) % nC).ravel())
    ci = np.arange(nC)[:, None]
    src.append(np.repeat(c0 + np.arange(nC), DEG))
    dst.append((b0 + (ci * DEG + j) % nB).ravel())
    off, ngh = O.csr_from_edges(n, np.concatenate(src), np.concatenate(dst))
    return off, ngh, s0, (b0, a0, c0, n)


def test_layered_graph_pages():
    off, ngh, s0, (b0, a0, c0, n) = layered_graph()
    pages = O.page_cuts(off, P)
    assert len(pages) == 14
    assert pages[0][0] == 0 and pages[0][1] == b0 - 1
    starts = [int(p[0]) for p in pages]
    assert starts[1] == b0 and starts[6] == a0 and starts[9] == c0
    lv = O.bfs(off, ngh, s0)
    assert (lv[a0:c0] == 1).all() and (lv[c0:] == 2).all() and (lv[b0:a0] == 3).all()
    assert O.lib().oracle_pool_size(3, n, 14, P) == 9


def run_mode(off, ngh, src, want, mode, extra=()):
    flags = ["-max_memory=3", f"-page_words={P}", "-build_pull=false",
             f"-page_predict={mode}", "-page_direct=0", *extra]
    with scaph_amd.BFS(*flags) as e:
        e.load_csr(off, ngh)
        assert not e.info()["resident"] and e.pool()["slots"] == 9
        if mode == "learned":
            L.model_for_graph(e, L.DEFAULT_MODEL)
        if mode == "exact":
            e.set_page_levels(want)
        st = e.run(src)
        np.testing.assert_array_equal(e.levels(), want, err_msg=mode)
        recs = e.records()
        say(mode, {k: st[k] for k in ("bytes_h2d", "page_execs", "page_execs_cached",
                                      "pages_full", "pages_compacted", "pages_prefetched")},
            [(r["level"], r["pages_full"], r["pages_prefetched"]) for r in recs])
        # a second run from the same source (the pool as the first left it)
        e.run(src)
        np.testing.assert_array_equal(e.levels(), want, err_msg=mode + " again")
        return st


def test_prediction_copies_fewer_bytes():
    off, ngh, s0, _ = layered_graph()
    want = O.bfs(off, ngh, s0)
    off_st = run_mode(off, ngh, s0, want, "off")
    lrn = run_mode(off, ngh, s0, want, "learned")
    ex = run_mode(off, ngh, s0, want, "exact")
    page = 16 * PER_PAGE * 4
    # plain policy: C evicts four B pages, which level 3 copies again
    assert off_st["pages_full"] == 9 and off_st["pages_prefetched"] == 0
    for st in (lrn, ex):
        assert st["bytes_h2d"] <= off_st["bytes_h2d"] - 3 * page
        assert st["page_execs_cached"] > off_st["page_execs_cached"]


@pytest.mark.parametrize("mode", ["learned", "exact"])
def test_prediction_keeps_levels_on_rmat_and_zipf(mode):
    """Bit-exact levels with the prediction on graphs of other shapes and
    small pools (look-ahead and whole-page copies do not change results)."""
    cases = [("rmat", O.rmat_csr(16, 16, seed=31), 2, "-threshold=16"),
             ("rmat-t3", O.rmat_csr(16, 16, seed=32), 3, "-threshold=3"),
             ("zipf", O.zipf_csr(120000, 1.2, 1.9, 20000, 5), 3, "-threshold=12")]
    for name, (off, ngh), mib, thr in cases:
        deg = np.diff(off.astype(np.int64))
        with scaph_amd.BFS(f"-max_memory={mib}", "-page_words=65536", "-build_pull=false", thr,
                           f"-page_predict={mode}") as e:
            e.load_csr(off, ngh)
            assert not e.info()["resident"]
            if mode == "learned":
                L.model_for_graph(e, L.DEFAULT_MODEL)
            for src in (int(np.argmax(deg)), int(np.nonzero(deg)[0][11])):
                want = O.bfs(off, ngh, src)
                if mode == "exact":
                    e.set_page_levels(want)
                st = e.run(src)
                say(name, mode, src, {k: st[k] for k in ("bytes_h2d", "pages_prefetched")})
                np.testing.assert_array_equal(e.levels(), want, err_msg=f"{name} {src}")
