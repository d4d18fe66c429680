"""CPU reference of the learned far-field model -- TEST INFRASTRUCTURE ONLY.

The model of the product library (8 -> 64 -> 64 -> 64 -> 2, GELU in tanh form, parameters
W1[64][8] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[2][64] b4[2]), its loss
L = (1/N) sum_k |y_k - t_k|^2 and the gradient by backpropagation, in FP64.  With
emulate=True every matrix-core operand is rounded to BF16 (round to nearest even) exactly
where the device rounds it (inputs, weights, activations, the backpropagated dZ), while
pre-activations, biases and bias gradients stay unrounded, so the device's results agree up
to FP32 accumulation order.  No reference counterpart: BatchLayout has no learned part.
The backpropagation is pinned by central finite differences in tests/test_oracle_mlp.py.
"""
import numpy as np

SHAPES = [("W1", (64, 8)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("W3", (64, 64)), ("b3", (64,)),
          ("W4", (2, 64)), ("b4", (2,))]
OFFSETS = {}
_o = 0
for _k, _s in SHAPES:
    OFFSETS[_k] = (_o, _s)
    _o += int(np.prod(_s))
NPARAM = _o  # 9026


def unpack(p):
    p = np.asarray(p, np.float64)
    return {k: p[o:o + int(np.prod(s))].reshape(s) for k, (o, s) in OFFSETS.items()}


def pack(d):
    out = np.zeros(NPARAM)
    for k, (o, s) in OFFSETS.items():
        out[o:o + int(np.prod(s))] = np.asarray(d[k]).ravel()
    return out


def bf16(x):
    """Round to BF16 (nearest even) via FP32, returned as float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)


_C = np.sqrt(2.0 / np.pi)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(_C * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    t = np.tanh(_C * (z + 0.044715 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * _C * (1.0 + 3 * 0.044715 * z * z)


def forward(p, X, emulate=True):
    q = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    P = unpack(p)
    H0 = q(X)
    Z1 = H0 @ q(P["W1"]).T + P["b1"]
    H1 = q(gelu(Z1))
    Z2 = H1 @ q(P["W2"]).T + P["b2"]
    H2 = q(gelu(Z2))
    Z3 = H2 @ q(P["W3"]).T + P["b3"]
    H3 = q(gelu(Z3))
    Y = H3 @ q(P["W4"]).T + P["b4"]
    return Y, (H0, Z1, H1, Z2, H2, Z3, H3)


def loss_grad(p, X, T, emulate=True):
    q = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    P = unpack(p)
    Y, (H0, Z1, H1, Z2, H2, Z3, H3) = forward(p, X, emulate)
    N = len(X)
    E = Y - np.asarray(T, np.float64)
    loss = (E ** 2).sum() / N
    g = {}
    dZ = 2.0 * E / N
    dZq = q(dZ)
    g["W4"], g["b4"] = dZq.T @ H3, dZ.sum(0)
    dH = dZq @ q(P["W4"])
    for k, Z, Hprev in (("3", Z3, H2), ("2", Z2, H1), ("1", Z1, H0)):
        dZ = dH * gelu_grad(Z)
        dZq = q(dZ)
        g["W" + k], g["b" + k] = dZq.T @ Hprev, dZ.sum(0)
        if k != "1":
            dH = dZq @ q(P["W" + k])
    return loss, pack(g)
