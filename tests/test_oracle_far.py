"""Pins of the oracle's physically consistent Barnes-Hut (orc_far_forces, orc_far_samples,
orc_far_layout; no reference counterpart) by brute force and by the multipole expansion of
the 2-D field d/|d|^2 = 1/conj(z): for a cell with mean c and offsets delta_p,
sum_p 1/conj(w + delta_p) = (m/conj(w)) (1 + sum conj(delta)^2/(m conj(w)^2) + O((delta/w)^3)),
so in the frame of u = w/|w| the corrections are a = S_uu - S_ww, b = -2 S_uw to second order."""
import numpy as np
import pytest


def _brute(xy, q):
    d = xy[None, :, :] - q[:, None, :]
    r2 = (d ** 2).sum(-1)
    with np.errstate(divide="ignore", invalid="ignore"):
        t = np.where(r2[..., None] > 0, d / r2[..., None], 0.0)
    return t.sum(1)


def test_theta_zero_is_the_exact_sum(oracle):
    rng = np.random.default_rng(0)
    xy = rng.random((700, 2))
    q = np.vstack([xy[:50], rng.random((50, 2))])
    got = oracle.far_forces(xy, 0.0, queries=q)
    ref = _brute(xy, q)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    ex = oracle.exact_repulsion(xy, np.arange(50))
    assert np.abs(got[:50] - ex).max() <= 1e-12 * np.abs(ex).max()


def test_single_far_cluster_is_its_monopole_plus_quadrupole(oracle):
    # one tight cluster far from the query: accepted as a whole (theta 0.5 opens the root,
    # whose side 5 is comparable to its distance)
    rng = np.random.default_rng(1)
    pts = np.array([5.0, 3.0]) + 0.01 * rng.standard_normal((64, 2))
    q = np.array([[0.0, 0.0]])
    xy = np.vstack([pts, [[0.0, 0.0]]])  # the query itself is a point (contributes 0)
    f = oracle.far_forces(xy, 0.5, queries=q)[0]
    c = pts.mean(0)
    w = c - q[0]
    mono = 64 * w / (w @ w)
    assert np.allclose(f, mono, rtol=1e-14, atol=0)  # the cluster is one accepted cell
    # and the monopole differs from the exact sum only at second order
    ex = _brute(pts, q)[0]
    assert np.linalg.norm(ex - mono) / np.linalg.norm(ex) < (0.01 * 4 / 5.8) ** 2


def test_samples_reconstruct_exact_cell_sums(oracle):
    rng = np.random.default_rng(2)
    xy = rng.random((3000, 2))
    perm = oracle.bh_tree(xy)["perm"]
    for qi in (0, 1234, 2999):
        feat, targ, cell, force = oracle.far_samples(xy, 1.0, xy[qi])
        assert len(feat) > 5
        v = xy[qi]
        mono_total = np.zeros(2)
        for f, (a, b), (lo, m) in zip(feat, targ, cell):
            P = xy[perm[lo:lo + m]]
            c = P.mean(0)
            w = c - v
            rho = np.linalg.norm(w)
            u = w / rho
            wv = np.array([-u[1], u[0]])
            exact_cell = _brute(P, v[None])[0]
            rec = (m / rho) * ((1 + a) * u + b * wv)
            assert np.allclose(rec, exact_cell, rtol=1e-12, atol=1e-12 * np.abs(exact_cell).max())
            # features: opening ratio below theta, second moments as defined
            assert 0 < f[0] < 1.0 and np.isclose(f[7], f[0] ** 2)
            D = (P - c)
            S = D.T @ D / (m * rho * rho)
            assert np.isclose(f[1] + f[3], np.trace(S), rtol=1e-12)
            assert np.isclose(f[1], u @ S @ u, rtol=1e-10, atol=1e-15)
            assert np.isclose(f[2], u @ S @ wv, rtol=1e-10, atol=1e-15)
            assert np.isclose(f[4], f[1] - f[3]) and np.isclose(f[5], np.log2(m) / 16) and np.isclose(f[6], 1 / m)
            # second-order multipole: a ~ S_uu - S_ww, b ~ -2 S_uw, error O(ratio^3)
            assert abs(a - (f[1] - f[3])) <= 0.2 * f[0] ** 3
            assert abs(b + 2 * f[2]) <= 0.2 * f[0] ** 3
            mono_total += m * w / (rho * rho)
        # the traversal's monopole force contains exactly these cells' monopoles
        assert np.linalg.norm(force - mono_total) <= np.linalg.norm(_brute(xy, v[None])[0]) * 2


def test_far_layout_theta_zero_first_step(oracle):
    # one vertex per batch, two vertices: closed form of one iteration
    rp = np.array([0, 1, 2], np.int32)
    ci = np.array([1, 0], np.int32)
    x0 = np.array([[0.0, 0.0], [3.0, 4.0]])
    out, e = oracle.far_layout(rp, ci, x0, 1, 1, 0.0)
    # vertex 0: attraction d*|d| = (3,4)*5, repulsion (3,4)/25 -> F = (15,20) - (0.12,0.16)
    F0 = np.array([15 - 0.12, 20 - 0.16])
    x1 = x0[0] + F0 / np.linalg.norm(F0)
    # vertex 1: attraction toward the moved vertex 0, repulsion from the iteration-start one
    d = x1 - x0[1]
    F1 = d * np.linalg.norm(d) - np.array([-0.12, -0.16])
    assert np.allclose(out[0], x1, rtol=0, atol=1e-15)
    assert np.allclose(out[1], x0[1] + F1 / np.linalg.norm(F1), rtol=0, atol=1e-14)
    assert np.isclose(e, F0 @ F0 + F1 @ F1, rtol=1e-14)
