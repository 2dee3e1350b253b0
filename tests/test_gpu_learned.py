"""GPU tests of the learned schedule: MLP kernels against the oracle, the
device decision against the host restatement, and levels that stay
bit-exact whatever the model decides."""
import ctypes

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
