"""Pins of the CPU reference of the learned far-field model (oracle/mlp_ref.py): the FP64
backpropagation against central finite differences, and the BF16 rounding against its
definition."""
import numpy as np

from oracle import mlp_ref as M


def _params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(M.NPARAM)
    for k, (o, s) in M.OFFSETS.items():
        n = int(np.prod(s))
        fan = s[1] if len(s) == 2 else 64
        p[o:o + n] = rng.standard_normal(n) / np.sqrt(fan) * (1.0 if k[0] == "W" else 0.1)
    return p


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = _params(1)
    X = rng.standard_normal((37, 8))
    T = rng.standard_normal((37, 2))
    loss, g = M.loss_grad(p, X, T, emulate=False)
    idx = np.concatenate([rng.choice(np.arange(o, o + int(np.prod(s))), min(6, int(np.prod(s))), replace=False)
                          for o, s in M.OFFSETS.values()])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss_grad(pp, X, T, emulate=False)[0] - M.loss_grad(pm, X, T, emulate=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    # a wrong derivative of the activation would show: the tanh-form derivative vs differences
    z = np.linspace(-5, 5, 101)
    fd = (M.gelu(z + 1e-6) - M.gelu(z - 1e-6)) / 2e-6
    assert np.abs(fd - M.gelu_grad(z)).max() < 1e-8


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, -2.5, 0.0])
    # 8 significant bits: 1 + 2^-8 is a tie -> even (1.0); 1 + 3*2^-8 -> 1 + 2^-6 (even), ...
    got = M.bf16(x)
    assert list(got) == [1.0, 1.0, 1.0 + 2 ** -6, 1.0, 1.0 + 2 ** -7, -2.5, 0.0]
    # emulated and exact forward agree to BF16 precision
    rng = np.random.default_rng(3)
    p = _params(2)
    X = rng.standard_normal((50, 8))
    y1, _ = M.forward(p, X, emulate=True)
    y2, _ = M.forward(p, X, emulate=False)
    assert np.abs(y1 - y2).max() < 0.05 * np.abs(y2).max()
