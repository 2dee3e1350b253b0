"""GPU tests of the learned schedule: MLP kernels against the oracle, the
device decision against the host restatement, and levels that stay
bit-exact whatever the model decides."""
import ctypes
import os

import numpy as np
import pytest

import scaph_amd
from scaph_amd import _lib
from scaph_amd import learn as L
from tests import _oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


def test_mlp_forward_backward_sgd_vs_oracle():
    say("mlp kernels")
    rng = np.random.default_rng(3)
    master = (rng.normal(0, 0.25, L.NPARAM)).astype(np.float32)
    pb = L.f32_to_bf16(master)
    pf = L.bf16_to_f32(pb)
    X = rng.normal(0, 1, (300, 16)).astype(np.float32)
    T = np.stack([rng.uniform(0, 1, 300), rng.uniform(-1, 1, 300)], 1).astype(np.float32)
    mlp = L.DeviceMLP(master)
    np.testing.assert_array_equal(mlp.params_bf16(), pb)
    # forward
    got = mlp.predict(X)
    want, _, _ = O.mlp_forward(pf, X)
    np.testing.assert_allclose(got, want, rtol=1e-4, atol=1e-4)
    # backward (fp32 accumulation in a different order: tolerance)
    dx = L.DevArray(X)
    dT = L.DevArray(T)
    loss = mlp.step(dx, dT, len(X), lr=0.0, mu=0.0, w1=0.7)
    g = mlp.grad.get(np.float32, L.NPARAM)
    l_want, g_want = O.mlp_loss_grad(pf, X, T, 0.7)
    assert abs(float(loss.get(np.float32, 1)[0]) - l_want) < 1e-4 * max(1, abs(l_want))
    np.testing.assert_allclose(g, g_want, rtol=2e-3, atol=2e-6)
    # sgd: master -= lr * (mu*mom + g); params = bf16(master)
    lr, mu = 0.1, 0.9
    mlp.step(dx, dT, len(X), lr=lr, mu=mu, w1=0.7)
    m2 = mlp.master.get(np.float32, L.NPARAM)
    g2 = mlp.grad.get(np.float32, L.NPARAM)
    # the lr=0 step left momentum = g
    want_m = master - lr * (mu * g + g2)
    np.testing.assert_allclose(m2, want_m, rtol=1e-6, atol=1e-6)
    np.testing.assert_array_equal(mlp.params_bf16(), L.f32_to_bf16(m2))


def test_training_reduces_loss_on_device():
    say("training")
    rng = np.random.default_rng(4)
    X = rng.normal(0, 1, (512, 16)).astype(np.float32)
    w = rng.normal(0, 1, 16)
    T = np.stack([(X @ w > 0).astype(np.float32), np.tanh(X[:, 0]).astype(np.float32)], 1)
    mlp, mean, scale, losses = L.train(X, T, steps=600, lr=0.05)
    assert losses[-1][1] < 0.5 * losses[0][1], losses
    pred = mlp.predict(((X - mean) / scale).astype(np.float32))
    assert ((pred[:, 0] > 0) == (T[:, 0] > 0.5)).mean() > 0.9


def test_learned_schedule_levels_bit_exact_and_explore():
    """Traces from explore(), a small model trained on them, then the learned
    schedule: levels equal the oracle's whatever the model chooses."""
    off, ngh = O.rmat_csr(15, 16, seed=11)
    deg = np.diff(off.astype(np.int64))
    srcs = [int(np.argmax(deg)), int(np.nonzero(deg)[0][5])]
    with scaph_amd.BFS("-max_memory=65536") as e:
        e.load_csr(off, ngh)
        recs = []
        for s in srcs:
            say("explore", s)
            r = e.explore(s, 1)
            assert len(r) >= 2
            # the features are the run-time features: level index and sizes
            assert r[0]["level"] == 0 and r[0]["n_f"] == 1
            assert abs(r[0]["x"][0] - 1.0) < 1e-6  # log2(1 + 1)
            recs += r
        X, T = L.features_labels(recs)
        mlp, mean, scale, _ = L.train(X, T, steps=200)
        e.set_model(mlp.params_bf16(), mean, scale)
        e.set_schedule("learned")
        for s in srcs:
            want = O.bfs(off, ngh, s)
            e.run(s)
            np.testing.assert_array_equal(e.levels(), want)
        # an adversarial model: a constant logit (always pull / always push)
        for bias in (30.0, -30.0):
            p = np.zeros(L.NPARAM, np.float32)
            p[-2] = bias
            e.set_model(L.f32_to_bf16(p), np.zeros(16, np.float32), np.ones(16, np.float32))
            e.set_schedule("learned")
            for s in srcs:
                e.run(s)
                np.testing.assert_array_equal(e.levels(), O.bfs(off, ngh, s))
        # a second, larger graph on the same engine: explore and the learned
        # schedule follow the new sizes
        off2, ngh2 = O.rmat_csr(17, 8, seed=12)
        e.load_csr(off2, ngh2)
        s2 = int(np.argmax(np.diff(off2.astype(np.int64))))
        say("explore on a larger graph", s2)
        r2 = e.explore(s2, 1)
        assert sum(1 for _ in r2) >= 2
        np.testing.assert_array_equal(e.levels(), O.bfs(off2, ngh2, s2))
        e.set_model(mlp.params_bf16(), mean, scale)
        e.set_schedule("learned")
        e.run(s2)
        np.testing.assert_array_equal(e.levels(), O.bfs(off2, ngh2, s2))


@pytest.mark.parametrize("page_words", [1 << 14, 128, 48])
def test_frontier_page_counts_bitmap_and_list(page_words):
    """Active vertices per page (the decision features) from a bitmap and
    from a vertex list, for page tables that fit LDS and that do not."""
    say("frontier page counts", page_words)
    n = 40000
    off, ngh = O.er_csr(n, 320000, 12)
    deg = np.diff(off).astype(np.uint32)
    pages = np.ascontiguousarray(O.page_cuts(off, page_words))
    rng = np.random.default_rng(page_words)
    stat = (rng.random(n) < 0.2).astype(np.uint8)
    want, _, _ = O.page_counts(pages, stat, deg)
    words = ((n + 63) // 64) * 2
    bits = np.zeros(words * 32, np.uint8)
    bits[:n] = stat
    bitmap = np.packbits(bits, bitorder="little").view(np.uint32)
    queue = rng.permutation(np.nonzero(stat)[0]).astype(np.uint32)
    d_pages = L.DevArray(pages)
    d_bm = L.DevArray(bitmap)
    d_q = L.DevArray(queue)
    d_len = L.DevArray(np.array([len(queue)], np.uint64))
    d_cnt = L.DevArray(np.full(len(pages), 7, np.uint32))
    hip = _lib.hip()
    for form in ("bitmap", "list"):
        _lib.check(hip.scaph_frontier_page_counts(d_bm.ptr if form == "bitmap" else None,
                                                  None if form == "bitmap" else d_q.ptr,
                                                  None if form == "bitmap" else d_len.ptr,
                                                  d_pages.ptr, len(pages), n, d_cnt.ptr, None))
        got = d_cnt.get(np.uint32, len(pages))
        np.testing.assert_array_equal(got, want, err_msg=f"{form} pages={len(pages)}")
        d_cnt.put(np.full(len(pages), 7, np.uint32))
    # the fused entry (counts + decision in one launch) equals the two-call
    # path, call after call (it leaves its workspace zero)
    params = L.f32_to_bf16(L.init_params(seed=3))
    norm = np.concatenate([rng.normal(0, 1, 16), rng.uniform(0.5, 2, 16)]).astype(np.float32)
    d_par, d_norm = L.DevArray(params), L.DevArray(norm)
    cnt = np.zeros(6, np.uint64)
    cnt[0], cnt[1], cnt[2] = len(queue), int(deg[queue].sum()), len(queue)
    d_c = L.DevArray(cnt)
    d_ws = L.DevArray(np.zeros(len(pages) + 1, np.uint32))
    d_out, d_out2 = L.DevArray(nbytes=128), L.DevArray(nbytes=128)
    m_u = int(deg.sum())
    for form in ("bitmap", "list", "bitmap", "list"):
        bm, q = (d_bm.ptr, None) if form == "bitmap" else (None, d_q.ptr)
        _lib.check(hip.scaph_frontier_page_counts(bm, q, d_c.ptr if q else None, d_pages.ptr,
                                                  len(pages), n, d_cnt.ptr, None))
        _lib.check(hip.scaph_mlp_decide(d_par.ptr, d_norm.ptr, d_cnt.ptr, d_pages.ptr, len(pages),
                                        d_c.ptr, m_u, 3, d_out.ptr, None))
        _lib.check(hip.scaph_mlp_decide_frontier(d_par.ptr, d_norm.ptr, bm, q, d_pages.ptr,
                                                 len(pages), n, d_c.ptr, m_u, 3, d_ws.ptr,
                                                 d_out2.ptr, None))
        a, b = d_out.get(np.float32, 32), d_out2.get(np.float32, 32)
        np.testing.assert_array_equal(a, b, err_msg=form)
        assert np.all(d_ws.get(np.uint32, len(pages)) == 0)
        hist = np.bincount(np.where(want > 0, 1 + np.clip(np.floor(
            np.log2(np.maximum(want, 1) / np.maximum(pages[:, 2], 1)) + 11), 0, 10), 0).astype(int),
            minlength=12) / len(pages)
        np.testing.assert_allclose(a[4:16], hist, rtol=1e-6)
        assert a[2] == np.count_nonzero(want)
        np.testing.assert_allclose(a[16:20], [np.log2(1 + len(queue)), np.log2(1 + cnt[1]),
                                              np.log2(1 + m_u - cnt[1]), 3.0], rtol=1e-5)


def test_trained_model_on_held_out_graph():
    """The committed model, on a graph outside its training set: at every
    level both directions are timed from the same state (explore); where
    one direction is clearly faster (by 2x or more) the model must pick it
    on at least 80% of those levels — a count of decisions, not a bound on
    wall time.  The per-level costs of the model's, the fixed rule's and the
    best choices are printed.  The learned schedule's levels equal the fixed
    schedule's."""
    p, mean0, scale, _ = L.load_model()
    pf = L.bf16_to_f32(p)
    with scaph_amd.BFS("-max_memory=65536") as e:
        e.gen_rmat_device(19, 16, 0.6, 0.16, 0.16, seed=777)
        n = e.info()["n"]
        mean = L.graph_mean(mean0, L.load_model_flags(), n, e.info()["m"])
        srcs = L.pick_sources(e, 3, seed=777)
        t_model = t_best = t_fixed = 0.0
        clear = right = 0
        for s in srcs:
            say("held-out explore", s)
            recs = e.explore(s, 2)
            X = np.array([r["x"] for r in recs], np.float32)
            y, _, _ = O.mlp_forward(pf, ((X - mean) / scale).astype(np.float32))
            prev_pull, prev_nf = False, 0
            for r, yy in zip(recs, y):
                tp, tq = r["t_push_ms"], r["t_pull_ms"]
                if max(tp, tq) >= 2.0 * min(tp, tq) and max(tp, tq) > 0.02:
                    clear += 1
                    right += int((yy[0] > 0) == (tq < tp))
                t_model += tq if yy[0] > 0 else tp
                t_best += min(tp, tq)
                m_u = 2.0 ** r["x"][2] - 1
                if not prev_pull:
                    pull = r["m_f"] > m_u / 14
                else:
                    pull = not (r["n_f"] < prev_nf and r["n_f"] < n / 24)
                t_fixed += tq if pull else tp
                prev_pull, prev_nf = tq < tp, r["n_f"]
            e.set_schedule("fixed")
            e.run(s)
            want = e.levels()
            e.set_model(p, mean, scale)
            e.set_schedule("learned")
            e.run(s)
            np.testing.assert_array_equal(e.levels(), want)
        say(f"step time: model {t_model:.3f} ms, fixed rule {t_fixed:.3f} ms, "
            f"per-level best {t_best:.3f} ms; clear levels {clear}, model right on {right}")
        assert clear >= 3
        assert right >= 0.8 * clear


def test_scalar_model_decision_fused_into_hand_off():
    """A model with zero W1 columns for the histogram decides inside the
    level hand-off (scaph_mlp_decide_publish): its record equals the
    two-kernel decision's logit and threshold and the oracle's forward pass
    on the four scalar features, the counters arrive in mapped host memory
    behind the flag, and the device counters are cleared."""
    hip = _lib.hip()
    rng = np.random.default_rng(11)
    master = L.init_params(seed=5, scalar=True)
    assert np.all(master[:64 * 16].reshape(64, 16)[:, 4:] == 0)
    params = L.f32_to_bf16(master)
    norm = np.concatenate([rng.normal(0, 1, 16), rng.uniform(0.5, 2, 16)]).astype(np.float32)
    d_par, d_norm = L.DevArray(params), L.DevArray(norm)
    words, decide_word, clear_words = 96, 64, 12
    m_u = 123456789
    pages = np.array([[0, 99, 100, 0], [100, 199, 100, 0]], np.uint32)
    d_pages = L.DevArray(pages)
    d_pc = L.DevArray(np.zeros(2, np.uint32))
    hp, dp = ctypes.c_void_p(), ctypes.c_void_p()
    _lib.check(hip.scaph_host_mapped_alloc(ctypes.byref(hp), ctypes.byref(dp), 4 * 512))
    try:
        host = np.ctypeslib.as_array(ctypes.cast(hp, ctypes.POINTER(ctypes.c_uint32)), (512,))
        host[:] = 0
        for seq, (n_next, m_next, level) in enumerate([(5, 40, 1), (700000, 9000000, 4),
                                                       (0, 0, 9)], start=1):
            src = np.zeros(words, np.uint32)
            cnt = src[:12].view(np.uint64)
            cnt[0], cnt[1], cnt[2] = n_next, m_next, 77
            d_src = L.DevArray(src)
            d_out = L.DevArray(nbytes=128)
            _lib.check(hip.scaph_mlp_decide(d_par.ptr, d_norm.ptr, d_pc.ptr, d_pages.ptr, 2,
                                            d_src.ptr, m_u, level, d_out.ptr, None))
            _lib.check(hip.scaph_mlp_decide_publish(
                d_par.ptr, d_norm.ptr, d_src.ptr, decide_word, m_u, level, words, clear_words,
                dp, ctypes.c_void_p(dp.value + 4 * 256), seq, None))
            _lib.check(hip.scaph_device_sync())
            assert host[256] == seq
            np.testing.assert_array_equal(host[:12], src[:12])
            assert np.all(d_src.get(np.uint32, words)[:clear_words] == 0)
            rec = host[decide_word:decide_word + 32].view(np.float32)
            two = d_out.get(np.float32, 32)
            np.testing.assert_array_equal(rec[:2], two[:2])
            x = np.zeros((1, 16), np.float32)
            x[0, :4] = [np.log2(1 + n_next), np.log2(1 + m_next), np.log2(1 + m_u - m_next),
                        level]
            np.testing.assert_allclose(rec[16:20], x[0, :4], rtol=1e-5)
            y, _, _ = O.mlp_forward(L.bf16_to_f32(params), ((x - norm[:16]) / norm[16:]))
            np.testing.assert_allclose(rec[:2], y[0], rtol=1e-4, atol=1e-4)
    finally:
        hip.scaph_host_pinned_free(hp)


def test_margin_falls_back_to_fixed_rule():
    """-learned_margin: the model decides a level only when its |pull logit|
    reaches the margin, otherwise the fixed alpha/beta rule does.  A margin
    above every logit gives the fixed schedule's directions level for level;
    margin 0 gives the sign of every recorded logit.  The first level has no
    decision yet and is the fixed rule's in both."""
    with scaph_amd.BFS("-max_memory=65536", "-learned_margin=1e9") as a, \
            scaph_amd.BFS("-max_memory=65536", "-learned_margin=0") as b:
        for e in (a, b):
            e.gen_rmat_device(18, 16, seed=41)
            L.model_for_graph(e)
        srcs = L.pick_sources(a, 3, seed=41)
        decided = 0
        for s in srcs:
            say("margin fallback", s)
            a.set_schedule("fixed")
            a.run(s)
            fixed = [r["dir"] for r in a.records()]
            want = a.levels()
            a.set_schedule("learned")
            a.run(s)
            assert [r["dir"] for r in a.records()] == fixed
            np.testing.assert_array_equal(a.levels(), want)
            b.set_schedule("learned")
            b.run(s)
            recs = b.records()
            assert recs[0]["dir"] == fixed[0] and recs[0]["logit"] == 0.0
            for r in recs[1:]:
                assert r["dir"] == int(r["logit"] > 0), (s, r["level"], r["logit"])
                decided += 1
            np.testing.assert_array_equal(b.levels(), want)
        assert decided >= 6


def test_model_flag_binds_graph_offsets():
    """-model=<file> (the command-line path) and learn.model_for_graph (the
    Python path) give the same per-level directions on the same graph: the
    engine folds the loaded graph's log2 n / log2 m into the offsets of a
    model whose features are shares.  bin/bfs with the model passes -check."""
    import subprocess
    with scaph_amd.BFS("-max_memory=65536", "-schedule=learned",
                       f"-model={L.DEFAULT_MODEL}") as a, \
            scaph_amd.BFS("-max_memory=65536") as b:
        for e in (a, b):
            e.gen_rmat_device(18, 16, seed=41)
        L.model_for_graph(b)
        b.set_schedule("learned")
        srcs = L.pick_sources(b, 3, seed=41)
        for s in srcs:
            say("model flag vs python path", s)
            a.run(s)
            da = [r["dir"] for r in a.records()]
            la = a.levels()
            b.run(s)
            assert da == [r["dir"] for r in b.records()]
            np.testing.assert_array_equal(la, b.levels())
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "scaph_amd",
                       "bin", "bfs")
    out = subprocess.run([exe, "-gen=rmat-device:scale=16,ef=16,seed=3", "-schedule=learned",
                          f"-model={L.DEFAULT_MODEL}", "-rounds=2", "-check", "-src=5"],
                         capture_output=True, text=True, timeout=100)
    assert out.returncode == 0 and "check passed!" in out.stdout, out.stdout + out.stderr
