"""Pins the oracle's per-edge-weight branch (an extension: the reference scales every attraction
by the global W only, sample/algorithms.cpp:759, :985, :912, :1055-1058).

The oracle multiplies W by w_edge[k] of the CSR entry k that meets the neighbour
(oracle/oracle.c: orc_minibatch, bh_iterations, orc_far_layout).  Three kinds of check:

  identity      w_edge filled with a constant c gives bit for bit the run with weight = c and no
                per-edge weights (W = weight * w = c either way), for every layout family;
  restatement   a numpy restatement written from the mathematics (attraction W w_k |d| d / W w_k d /
                d log2(1 + W w_k |d|) toward each CSR neighbour with live positions; repulsion
                -d/|d|^2 or unit vectors from the positions the reference uses; sequential batches;
                unit-step move) agrees with the oracle on small graphs with random, asymmetric
                per-entry weights to FP64 rounding;
  mutation      the same restatement with the weights indexed by column id instead of CSR entry
                (a plausible mistake), or with the weights of the two directions of an edge
                swapped, disagrees with the oracle far beyond that rounding.

Two-vertex closed forms with different weights on the two directions of the edge complete it.
"""
import numpy as np
import pytest

from batchlayoutcode_amd import graphs


def _small_graph(rows=6, cols=7, seed=3):
    n, e = graphs.grid(rows, cols)
    rng = np.random.default_rng(seed)
    extra = rng.integers(0, n, size=(12, 2))  # a few long edges, so batches see far neighbours
    e = graphs._canon(np.concatenate([e, extra]))
    rp, ci = graphs.csr_from_edges(n, e)
    w = rng.uniform(0.5, 2.0, len(ci)).astype(np.float32)  # asymmetric: one weight per CSR entry
    xy = rng.random((n, 2))
    return rp, ci, w, xy


def _attract(law, W, d):
    r = np.sqrt(d[0] * d[0] + d[1] * d[1])
    if law == 2:
        return d * (W * r)
    if law == 8:
        return d * W
    return d * np.log2(1.0 + W * r)


def _np_minibatch(rp, ci, w, xy, iters, B, law, weight=1.0, windex=None):
    """Algo 2/3/8 from the mathematics: neighbours (the sorted, duplicate-free row) attract,
    every other vertex at |d| > 0 repels with -d/|d|^2; batches of B consecutive ids move
    one after another; inside a batch every force uses the positions before the batch."""
    n = len(rp) - 1
    xy = xy.copy()
    step, decay = 1.0, (0.999 if law == 2 else 0.99)
    E = 0.0
    widx = np.arange(len(ci)) if windex is None else windex
    for _ in range(iters):
        E = 0.0
        for b0 in range(0, n, B):
            F = np.zeros((min(B, n - b0), 2))
            for t, i in enumerate(range(b0, min(b0 + B, n))):
                nb = set(ci[rp[i]:rp[i + 1]].tolist())
                d = xy - xy[i]
                r2 = (d * d).sum(1)
                rep = np.array([j not in nb and r2[j] > 0 for j in range(n)])
                f = -(d[rep] / r2[rep][:, None]).sum(0)
                for k in range(rp[i], rp[i + 1]):
                    f = f + _attract(law, weight * float(w[widx[k]]), d[ci[k]])
                F[t] = f
            for t, i in enumerate(range(b0, min(b0 + B, n))):
                m = np.sqrt((F[t] ** 2).sum())
                xy[i] = xy[i] + F[t] / m * step
                E += (F[t] ** 2).sum()
        step *= decay
    return xy, E


def _np_algo4_like(rp, ci, w, xy, iters, B, rep_of, weight=1.0, windex=None):
    """The Barnes-Hut iteration shape (algorithms.cpp:1020-1083): repulsion rep_of(start positions)
    fixed for the iteration, attraction W w_k |d| d over every CSR entry with live positions,
    sequential batches, STEP *= 0.999."""
    n = len(rp) - 1
    xy = xy.copy()
    step, E = 1.0, 0.0
    widx = np.arange(len(ci)) if windex is None else windex
    for _ in range(iters):
        E = 0.0
        R = rep_of(xy.copy())
        for b0 in range(0, n, B):
            idx = range(b0, min(b0 + B, n))
            F = []
            for i in idx:
                f = np.zeros(2)
                for k in range(rp[i], rp[i + 1]):
                    f = f + _attract(2, weight * float(w[widx[k]]), xy[ci[k]] - xy[i])
                F.append(f - R[i])
            for i, f in zip(idx, F):
                xy[i] = xy[i] + f / np.sqrt((f ** 2).sum()) * step
                E += (f ** 2).sum()
        step *= 0.999
    return xy, E


def _unit_rep(x):
    d = x[None, :, :] - x[:, None, :]
    r = np.sqrt((d ** 2).sum(-1))
    np.fill_diagonal(r, np.inf)
    return (d / r[..., None]).sum(1)


def _exact_rep(x):
    d = x[None, :, :] - x[:, None, :]
    r2 = (d ** 2).sum(-1)
    np.fill_diagonal(r2, np.inf)
    return (d / r2[..., None]).sum(1)


def _colid_index(rp, ci):
    return ci.copy()  # the mistake: w[colid] instead of w[entry]


def _swap_directions(rp, ci):
    """Index of the reverse entry (j -> i) for every entry (i -> j)."""
    n = len(rp) - 1
    pos = {}
    for i in range(n):
        for k in range(rp[i], rp[i + 1]):
            pos[(i, int(ci[k]))] = k
    return np.array([pos[(int(ci[k]), i)] for i in range(n) for k in range(rp[i], rp[i + 1])])


# ---- identity: constant per-edge weights == the global W --------------------------------

@pytest.mark.parametrize("c", [2.0, 0.75])
def test_constant_edge_weight_equals_global_weight(oracle, c):
    rp, ci, _, xy = _small_graph()
    wc = np.full(len(ci), c, np.float32)
    for law in (2, 3, 8):
        a, ea, _ = oracle.minibatch(rp, ci, xy, 3, 8, law=law, weight=c)
        b, eb, _ = oracle.minibatch(rp, ci, xy, 3, 8, law=law, weight=1.0, w_edge=wc)
        assert np.array_equal(a, b) and ea == eb, law
    a, ea = oracle.bh_layout(rp, ci, xy, 3, 8, theta=1.2, weight=c)
    b, eb = oracle.bh_layout(rp, ci, xy, 3, 8, theta=1.2, weight=1.0, w_edge=wc)
    assert np.array_equal(a, b) and ea == eb
    a, ea = oracle.far_layout(rp, ci, xy, 3, 8, 0.8, weight=c)
    b, eb = oracle.far_layout(rp, ci, xy, 3, 8, 0.8, weight=1.0, w_edge=wc)
    assert np.array_equal(a, b) and ea == eb


# ---- restatement and mutations ----------------------------------------------------------

@pytest.mark.parametrize("law", [2, 3, 8])
def test_weighted_minibatch_matches_restatement(oracle, law):
    rp, ci, w, xy = _small_graph()
    B, iters, W = 8, 2, 1.5
    got, eg, _ = oracle.minibatch(rp, ci, xy, iters, B, law=law, weight=W, w_edge=w)
    ref, er = _np_minibatch(rp, ci, w, xy, iters, B, law, weight=W)
    diam = float(np.ptp(ref, axis=0).max())
    assert np.abs(got - ref).max() <= 1e-11 * diam and abs(eg - er) <= 1e-11 * er
    for bad in (_colid_index(rp, ci), _swap_directions(rp, ci)):
        mut, _ = _np_minibatch(rp, ci, w, xy, iters, B, law, weight=W, windex=bad)
        assert np.abs(got - mut).max() > 1e-6 * diam


def test_weighted_bh_layout_matches_restatement(oracle):
    # theta -> 0 opens every cell, so calcRepForce is the brute-force sum of unit vectors
    rp, ci, w, xy = _small_graph(seed=4)
    got, eg = oracle.bh_layout(rp, ci, xy, 2, 8, theta=1e-12, weight=1.25, w_edge=w)
    ref, er = _np_algo4_like(rp, ci, w, xy, 2, 8, _unit_rep, weight=1.25)
    diam = float(np.ptp(ref, axis=0).max())
    assert np.abs(got - ref).max() <= 1e-11 * diam and abs(eg - er) <= 1e-11 * er
    for bad in (_colid_index(rp, ci), _swap_directions(rp, ci)):
        mut, _ = _np_algo4_like(rp, ci, w, xy, 2, 8, _unit_rep, weight=1.25, windex=bad)
        assert np.abs(got - mut).max() > 1e-6 * diam


def test_weighted_far_layout_matches_restatement(oracle):
    # theta = 0 is the exact d/|d|^2 sum
    rp, ci, w, xy = _small_graph(seed=5)
    got, eg = oracle.far_layout(rp, ci, xy, 2, 8, 0.0, weight=0.8, w_edge=w)
    ref, er = _np_algo4_like(rp, ci, w, xy, 2, 8, _exact_rep, weight=0.8)
    diam = float(np.ptp(ref, axis=0).max())
    assert np.abs(got - ref).max() <= 1e-11 * diam and abs(eg - er) <= 1e-11 * er
    for bad in (_colid_index(rp, ci), _swap_directions(rp, ci)):
        mut, _ = _np_algo4_like(rp, ci, w, xy, 2, 8, _exact_rep, weight=0.8, windex=bad)
        assert np.abs(got - mut).max() > 1e-6 * diam


# ---- two-vertex closed forms --------------------------------------------------------------

def test_weighted_two_vertex_closed_form(oracle):
    # one edge stored as entries 0 (0 -> 1, weight a) and 1 (1 -> 0, weight b); B = 1.  Vertex 0
    # feels W a |d| d toward 1 and moves one unit; then vertex 1 feels W b |d'| d' toward the
    # moved vertex 0.  No repulsion between neighbours.
    rp = np.array([0, 1, 2], np.int32)
    ci = np.array([1, 0], np.int32)
    a, b, W = 0.5, 3.0, 2.0
    w = np.array([a, b], np.float32)
    xy0 = np.array([[0.0, 0.0], [3.0, 4.0]])
    f0 = np.array([3.0, 4.0]) * (W * a * 5.0)
    x0 = xy0[0] + f0 / np.linalg.norm(f0)
    d1 = x0 - xy0[1]
    f1 = d1 * (W * b * np.linalg.norm(d1))
    x1 = xy0[1] + f1 / np.linalg.norm(f1)
    E = (f0 ** 2).sum() + (f1 ** 2).sum()
    xy, e, _ = oracle.minibatch(rp, ci, xy0, 1, 1, weight=W, w_edge=w)
    assert np.allclose(xy, [x0, x1], rtol=0, atol=1e-15) and np.isclose(e, E, rtol=1e-15)
    _, e_swapped, _ = oracle.minibatch(rp, ci, xy0, 1, 1, weight=W, w_edge=w[::-1].copy())
    assert abs(e_swapped - E) > 1e-3 * E
    # the algo-4 and far-field iterations with two vertices: the only other vertex is a
    # neighbour, so repulsion and attraction both act, from the start and live positions
    for run, rep in ((lambda: oracle.bh_layout(rp, ci, xy0, 1, 1, theta=1.2, weight=W, w_edge=w), 1),
                     (lambda: oracle.far_layout(rp, ci, xy0, 1, 1, 0.0, weight=W, w_edge=w), 2)):
        got, eg = run()
        d = xy0[1] - xy0[0]
        r0 = d / np.linalg.norm(d) if rep == 1 else d / (d ** 2).sum()  # repulsion on vertex 0, start positions
        g0 = d * (W * a * 5.0) - r0
        y0 = xy0[0] + g0 / np.linalg.norm(g0)
        r1 = -r0  # on vertex 1, from the start positions too
        e1 = y0 - xy0[1]
        g1 = e1 * (W * b * np.linalg.norm(e1)) - r1
        y1 = xy0[1] + g1 / np.linalg.norm(g1)
        assert np.allclose(got, [y0, y1], rtol=0, atol=1e-14)
        assert np.isclose(eg, (g0 ** 2).sum() + (g1 ** 2).sum(), rtol=1e-14)
        # two vertices move along their edge whatever the weights, so the energy is what
        # tells the directions' weights apart: swapped, the first force alone changes it
        h0 = d * (W * b * 5.0) - r0
        assert abs((h0 ** 2).sum() - (g0 ** 2).sum()) > 1e-3 * eg
