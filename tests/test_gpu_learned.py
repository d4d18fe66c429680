"""Learned far field (no reference counterpart): the device's BF16-MFMA forward, backward and
Adam kernels against the CPU reference of oracle/mlp_ref.py, the training-sample generator
against the FP64 oracle, and what the trained model gains over the plain monopole."""
import numpy as np
import pytest

from oracle import mlp_ref as M

pytestmark = pytest.mark.gpu


def _data(rng, n):
    X = np.abs(rng.standard_normal((n, 8))).astype(np.float32) * 0.5
    T = (0.3 * np.sin(3 * X[:, :2]) + 0.1 * X[:, 2:4] ** 2).astype(np.float32)
    return X, T


def test_forward_matches_reference(bl):
    rng = np.random.default_rng(0)
    p = bl.far_model_init(7)
    p[M.OFFSETS["W4"][0]:M.OFFSETS["W4"][0] + 128] *= 10  # outputs of order one
    X, _ = _data(rng, 1000)  # not a multiple of 64: partial last tile
    print("[case] mlp forward", flush=True)
    y = bl.far_model_eval(p, X)
    ref, _ = M.forward(p, X, emulate=True)
    exact, _ = M.forward(p, X, emulate=False)
    scale = np.abs(exact).max()
    assert np.abs(y - ref).max() <= 2e-3 * scale
    assert np.abs(y - exact).max() <= 3e-2 * scale


def test_gradient_matches_reference(bl):
    rng = np.random.default_rng(1)
    p = bl.far_model_init(3)
    p[M.OFFSETS["W4"][0]:M.OFFSETS["W4"][0] + 128] *= 10
    X, T = _data(rng, 777)
    print("[case] mlp gradient", flush=True)
    loss, g = bl.far_model_grad(p, X, T)
    rl, rg = M.loss_grad(p, X, T, emulate=True)
    el, eg = M.loss_grad(p, X, T, emulate=False)
    assert abs(loss - rl) <= 1e-3 * rl
    for k, (o, s) in M.OFFSETS.items():
        n = int(np.prod(s))
        a, b, c = g[o:o + n], rg[o:o + n], eg[o:o + n]
        assert np.linalg.norm(a - b) <= 1e-2 * np.linalg.norm(b), k
        assert np.linalg.norm(a - c) <= 5e-2 * np.linalg.norm(c), k


def test_training_lowers_the_loss_and_is_deterministic(bl):
    rng = np.random.default_rng(2)
    X, T = _data(rng, 20000)
    p0 = bl.far_model_init(11)
    print("[case] mlp training", flush=True)
    p1, losses = bl.far_model_train(p0, X, T, epochs=8, batch=1024, lr=3e-3, seed=5)
    p2, _ = bl.far_model_train(p0, X, T, epochs=8, batch=1024, lr=3e-3, seed=5)
    assert np.array_equal(p1, p2)
    assert losses[-1] < 0.2 * losses[0]
    l_host, _ = M.loss_grad(p1, X, T, emulate=True)
    assert abs(l_host - losses[-1]) <= 0.5 * losses[-1]


def test_samples_match_oracle(bl, oracle):
    rng = np.random.default_rng(4)
    xy = rng.random((4000, 2))
    q = xy[[3, 1000, 3999]]
    print("[case] far samples", flush=True)
    feat, targ = bl.far_samples(xy, 1.0, q, cap=1024)
    rows = 0
    for k in range(len(q)):
        f_ref, t_ref, _, _ = oracle.far_samples(xy, 1.0, q[k])
        f = feat[rows:rows + len(f_ref)]
        t = targ[rows:rows + len(f_ref)]
        rows += len(f_ref)
        assert np.allclose(f, f_ref, rtol=1e-6, atol=1e-7)
        assert np.allclose(t, t_ref, rtol=1e-4, atol=1e-6)
    assert rows == len(feat)


def _train_model(bl, rng, theta):
    feats, targs = [], []
    for g in range(6):
        n = 6000
        if g % 2 == 0:
            xy = rng.random((n, 2))
        else:
            c = rng.random((10, 2))
            xy = c[rng.integers(0, 10, n)] + 0.04 * rng.standard_normal((n, 2))
        q = xy[rng.choice(n, 300, replace=False)]
        f, t = bl.far_samples(xy, theta, q, cap=512)
        feats.append(f)
        targs.append(t / f[:, 7:8])  # the model predicts (a, b) / (side/rho)^2
    F = np.concatenate(feats)
    T = np.concatenate(targs)
    p, losses = bl.far_model_train(bl.far_model_init(1), F, T, epochs=30, batch=4096, lr=2e-3, seed=3)
    return p, losses


def test_learned_far_field_beats_the_monopole(bl):
    rng = np.random.default_rng(6)
    theta = 1.0
    print("[case] learned far field ablation", flush=True)
    p, losses = _train_model(bl, rng, theta)
    assert losses[-1] < losses[0]
    # held-out point set
    c = rng.random((10, 2))
    xy = np.vstack([rng.random((3000, 2)), c[rng.integers(0, 10, 3000)] + 0.04 * rng.standard_normal((3000, 2))])
    q = xy[rng.choice(len(xy), 500, replace=False)]
    ex = bl.far_forces(xy, 0.0, mode=bl.FAR_EXACT, queries=q)
    err = {}
    for name, mode in (("mono", bl.FAR_MONO), ("quad", bl.FAR_QUAD), ("learned", bl.FAR_LEARNED)):
        f = bl.far_forces(xy, theta, mode=mode, queries=q, model=p)
        err[name] = np.median(np.linalg.norm(f - ex, axis=1) / np.linalg.norm(ex, axis=1))
    print("median relative force error", err, flush=True)
    assert err["learned"] < 0.7 * err["mono"]
    assert err["quad"] < err["mono"]
