"""Physically consistent Barnes-Hut (bl_far_forces, algos 40/41; no reference counterpart):
the mass-weighted monopole traversal and the exact sum against the FP64 oracle, and the
layouts built on them against the oracle's layout loop."""
import numpy as np
import pytest

from batchlayoutcode_amd import graphs
from conftest import diameter

pytestmark = pytest.mark.gpu


def _abs_scale(xy, q):
    """sum_j 1/|x_j - q|: the size of the terms whose rounding the tolerance covers."""
    d = np.sqrt(((xy[None, :, :] - q[:, None, :]) ** 2).sum(-1))
    with np.errstate(divide="ignore"):
        inv = np.where(d > 0, 1.0 / d, 0.0)
    return inv.sum(1)


def _clustered(rng, n):
    c = rng.random((8, 2))
    k = rng.integers(0, 8, n)
    return c[k] + 0.03 * rng.standard_normal((n, 2))


@pytest.mark.parametrize("kind", ["uniform", "clustered"])
@pytest.mark.parametrize("theta", [0.3, 0.8, 1.2])
def test_monopole_matches_oracle(bl, oracle, kind, theta):
    rng = np.random.default_rng(5)
    xy = rng.random((3000, 2)) if kind == "uniform" else _clustered(rng, 3000)
    print(f"[case] far monopole {kind} theta {theta}", flush=True)
    got = bl.far_forces(xy, theta)
    ref = oracle.far_forces(xy, theta)
    scale = _abs_scale(xy, xy)
    assert np.all(np.abs(got - ref).max(1) <= 1e-12 * scale)


def test_exact_and_theta_zero(bl, oracle):
    rng = np.random.default_rng(6)
    xy = _clustered(rng, 2000)
    q = rng.random((300, 2))
    print("[case] far exact", flush=True)
    ex = bl.far_forces(xy, 0.0, mode=bl.FAR_EXACT, queries=q)
    mono0 = bl.far_forces(xy, 0.0, queries=q)
    ref = oracle.far_forces(xy, 0.0, queries=q)
    scale = _abs_scale(xy, q)
    assert np.all(np.abs(ex - ref).max(1) <= 1e-12 * scale)
    assert np.all(np.abs(mono0 - ref).max(1) <= 1e-12 * scale)
    # the monopole error shrinks with theta (second-order in side/distance)
    errs = []
    for th in (1.2, 0.6, 0.3):
        f = bl.far_forces(xy, th, queries=q)
        errs.append(np.median(np.linalg.norm(f - ex, axis=1) / np.linalg.norm(ex, axis=1)))
    assert errs[0] > errs[1] > errs[2] and errs[0] < 0.05


@pytest.mark.parametrize("algo,theta", [(40, 0.8), (41, 0.0)])
def test_far_layout_matches_oracle(bl, oracle, algo, theta):
    rp, ci = graphs.csr_from_edges(*graphs.grid(20, 20))
    x0 = np.random.default_rng(8).random((400, 2))
    print(f"[case] far layout algo {algo}", flush=True)
    with bl.Graph(rp, ci) as g:
        got, st = g.run(x0, algo=algo, iters=3, batch=32, bht=theta)
    ref, e_ref = oracle.far_layout(rp, ci, x0, 3, 32, theta)
    assert np.abs(got - ref).max() <= 1e-9 * diameter(ref)
    assert abs(st.energy - e_ref) <= 1e-9 * e_ref
