"""Pins the CPU oracle (oracle/oracle.c) -- the checker every GPU parity test relies on.

What pins each oracle function:
  orc_mtx_read       CSR the reference itself built from every fixture input (csr_*.npz),
                     including the symmetric/general/diagonal/duplicate edge cases
  orc_minibatch      coordinates and energies the strict-IEEE reference build wrote after
                     k iterations (traj_a2_*, traj_a3_*, traj_a8_*, converged traj_a21_*):
                     bit-for-bit; a two-vertex closed form; the STEP schedule
  orc_bh_layout      traj_a4_* bit-for-bit; orc_hybrid_layout: traj_a6_* bit-for-bit
  orc_morton         scales the strict reference computed (bh_*.npz); the comparator of
                     MortonCode.h:83-97 restated literally and compared on random pairs
  orc_bh_tree        the reference's own tree dump (pre-order, full precision), bit-for-bit
  orc_bh_forces      calcRepForce values of the strict reference, bit-for-bit; the theta->0
                     limit equals the brute-force sum of unit vectors; theta->inf accepts
                     the root
  orc_exact_repulsion brute force in numpy
Mutation checks show the fixtures catch plausible mistakes (x bit above y bit, a
mass-weighted centroid, repulsion on neighbours, 0.99 instead of 0.999).
"""
import numpy as np
import pytest

from conftest import diameter, golden, golden_files, golden_input


@pytest.mark.parametrize("fx", golden_files("csr_*.npz"))
def test_mtx_csr_matches_reference(oracle, fx):
    d = golden(fx)
    rp, ci = oracle.mtx_read(golden_input(fx[4:-4] + ".mtx"))
    assert np.array_equal(rp, d["rowptr"]) and np.array_equal(ci, d["colids"])


def _run_oracle(oracle, d, k):
    rp, ci = oracle.mtx_read(golden_input(str(d["input"])))
    algo, B, W = int(d["algo"]), int(d["batch"]), float(d["weight"])
    if algo in (2, 3, 8):
        xy, e, _ = oracle.minibatch(rp, ci, d["init"], int(k), B, law=algo, weight=W)
    elif algo == 21:
        xy, e, _ = oracle.minibatch(rp, ci, d["init"], 0, B, law=2, weight=W, converged=True, engt=float(d["engt"]))
    elif algo == 4:
        xy, e = oracle.bh_layout(rp, ci, d["init"], int(k), B, theta=float(d["bht"]), weight=W)
    else:
        xy, e = oracle.hybrid_layout(rp, ci, d["init"], int(k), B, weight=W)
    return xy, e


@pytest.mark.parametrize("fx", golden_files("traj_*.npz"))
def test_trajectory_bit_exact_vs_strict_reference(oracle, fx):
    d = golden(fx)
    ks = list(d["ks"])
    if "grid64" in fx:
        ks = ks[:4]  # keep the CPU suite short; later iterations are covered on small graphs
    for kk, k in enumerate(ks):
        xy, e = _run_oracle(oracle, d, k)
        assert np.array_equal(xy, d["xy_strict"][kk]), (fx, k)
        assert e == d["E_strict"][kk], (fx, k)
        # and within FP64 rounding noise of the fast-math build for the first iterations
        if k <= 3:
            assert np.abs(xy - d["xy"][kk]).max() <= 1e-8 * diameter(d["init"])


@pytest.mark.parametrize("fx", golden_files("bh_*.npz"))
def test_bh_structures_vs_reference(oracle, fx):
    d = golden(fx)
    pts = d["pts"]
    xs, ys, mm = oracle.morton(pts)
    assert mm == d["maxmin"]
    assert np.array_equal(xs, d["scales_strict"][:, 0]) and np.array_equal(ys, d["scales_strict"][:, 1])
    # the fast-math build may differ only by one unit on an extreme coordinate
    dfx = np.abs(xs.astype(np.int64) - d["scales"][:, 0]) + np.abs(ys.astype(np.int64) - d["scales"][:, 1])
    assert dfx.max() <= 1 and (dfx > 0).sum() <= 1
    t = oracle.bh_tree(pts)
    assert np.array_equal(t["perm"], d["perm"]) and np.array_equal(t["perm"], d["perm_strict"])
    assert np.array_equal(t["parent"], d["node_parent"]) and np.array_equal(t["quad"], d["node_quad"])
    assert np.array_equal(t["diam"], d["node_diam"])
    assert np.array_equal(t["cx"], d["node_cx"]) and np.array_equal(t["cy"], d["node_cy"])
    leaves = sorted((v, q, -c - 2) for v in range(len(t["parent"])) for q, c in enumerate(t["child"][v]) if c <= -2)
    assert leaves == sorted(map(tuple, d["leaf"].tolist()))
    F = oracle.bh_forces(pts, float(d["theta"]))
    assert np.array_equal(F, d["force_strict"])


# ---- independent pins --------------------------------------------------------------

def _ref_comparator(a, b):
    """MortonCode::operator() (MortonCode.h:83-97) restated literally: a < b."""
    A = [a[1], a[0]]  # {yscale, xscale}
    Bv = [b[1], b[0]]
    j, x = 0, 0
    for k in range(2):
        y = A[k] ^ Bv[k]
        if x < y and x < (x ^ y):  # less_msb(x, y)
            j, x = k, y
    return A[j] < Bv[j]


def _interleave(xs, ys, y_high=True):
    key = 0
    for b in range(31, -1, -1):
        hi, lo = ((ys >> b) & 1, (xs >> b) & 1) if y_high else ((xs >> b) & 1, (ys >> b) & 1)
        key = (key << 2) | (hi << 1) | lo
    return key


def test_morton_comparator_equals_interleaved_key():
    rng = np.random.default_rng(5)
    for _ in range(4000):
        a = rng.integers(0, 2**32, size=2, dtype=np.uint64).tolist()
        b = rng.integers(0, 2**32, size=2, dtype=np.uint64).tolist()
        if rng.random() < 0.5:  # share high bits to exercise deep levels
            sh = int(rng.integers(1, 32))
            b[0] = (a[0] >> sh << sh) | (b[0] & ((1 << sh) - 1))
            b[1] = (a[1] >> sh << sh) | (b[1] & ((1 << sh) - 1))
        ka, kb = _interleave(a[0], a[1]), _interleave(b[0], b[1])
        assert _ref_comparator(a, b) == (ka < kb)


def test_morton_fixture_rejects_x_over_y(oracle):
    d = golden("bh_u1000_t12.npz")
    xs, ys = d["scales_strict"][:, 0].astype(object), d["scales_strict"][:, 1].astype(object)
    wrong = np.argsort([_interleave(int(x), int(y), y_high=False) for x, y in zip(xs, ys)], kind="stable")
    right = np.argsort([_interleave(int(x), int(y)) for x, y in zip(xs, ys)], kind="stable")
    assert np.array_equal(right, d["perm"]) and not np.array_equal(wrong, d["perm"])


def test_bh_theta_zero_is_brute_force_unit_vectors(oracle):
    pts = np.random.default_rng(7).random((300, 2))
    F = oracle.bh_forces(pts, 1e-12)
    d = pts[None, :, :] - pts[:, None, :]
    r = np.sqrt((d ** 2).sum(-1))
    np.fill_diagonal(r, np.inf)
    ref = (d / r[..., None]).sum(1)
    assert np.abs(F - ref).max() < 1e-10


def test_bh_huge_theta_accepts_root(oracle):
    pts = np.random.default_rng(8).random((200, 2))
    t = oracle.bh_tree(pts)
    c = np.array([t["cx"][0], t["cy"][0]])
    F = oracle.bh_forces(pts, 1e30)
    u = c[None, :] - pts
    u /= np.sqrt((u ** 2).sum(1))[:, None]
    assert np.abs(F - u).max() < 1e-14


def test_bh_fixture_rejects_weighted_centroid(oracle):
    # a mass-weighted centroid (count of points below) would differ from the reference's
    # unweighted mean of children at some node
    d = golden("bh_u1000_t12.npz")
    pts = d["pts"]
    t = oracle.bh_tree(pts)
    m = len(t["parent"])
    cnt = np.zeros(m)
    wx = np.zeros(m)
    for v in range(m - 1, -1, -1):
        for c in t["child"][v]:
            if c >= 0:
                cnt[v] += cnt[c]
                wx[v] += wx[c]
            elif c <= -2:
                cnt[v] += 1
                wx[v] += pts[-c - 2, 0]
    assert not np.allclose(wx / cnt, d["node_cx"], rtol=0, atol=1e-12)


def test_two_vertex_closed_form(oracle):
    # one edge, B=1: vertex 0 feels attraction W|d|d toward 1 (no repulsion on a neighbour)
    rp = np.array([0, 1, 2], np.int32)
    ci = np.array([1, 0], np.int32)
    xy0 = np.array([[0.0, 0.0], [3.0, 4.0]])
    xy, e, _ = oracle.minibatch(rp, ci, xy0, 1, 1, weight=2.0)
    f0 = np.array([3.0, 4.0]) * 2.0 * 5.0
    x0 = xy0[0] + f0 / np.linalg.norm(f0)
    d1 = x0 - xy0[1]
    f1 = d1 * 2.0 * np.linalg.norm(d1)
    x1 = xy0[1] + f1 / np.linalg.norm(f1)
    assert np.allclose(xy, [x0, x1], rtol=0, atol=1e-15)
    assert np.isclose(e, (f0 ** 2).sum() + (f1 ** 2).sum(), rtol=1e-15)


def test_step_schedule_and_decay_pinned(oracle):
    d = golden("traj_a2_grid8_b16.npz")
    rp, ci = oracle.mtx_read(golden_input("grid8x8.mtx"))
    # 3 iterations == 1 iteration, then 2 iterations from STEP = 0.999
    a, ea, _ = oracle.minibatch(rp, ci, d["init"], 1, 16)
    b, eb, _ = oracle.minibatch(rp, ci, a, 2, 16, step0=0.999)
    assert np.array_equal(b, d["xy_strict"][2]) and eb == d["E_strict"][2]
    # the wrong decay is caught by the fixture
    w, _, _ = oracle.minibatch(rp, ci, d["init"], 3, 16, decay=0.99)
    assert not np.allclose(w, d["xy_strict"][2], atol=1e-9)


def test_neighbour_repulsion_mutation_is_caught(oracle):
    # adding repulsion to neighbours too changes the iteration-1 layout the reference wrote
    d = golden("traj_a2_test_b256.npz")
    rp, ci = oracle.mtx_read(golden_input("test.mtx"))
    n = len(rp) - 1
    xy = d["init"].copy()
    F = np.zeros((n, 2))
    for i in range(n):
        dd = xy - xy[i]
        r2 = (dd ** 2).sum(1)
        r2[i] = np.inf
        F[i] = -(dd / r2[:, None]).sum(0)
        for j in ci[rp[i]:rp[i + 1]]:
            F[i] += (xy[j] - xy[i]) * np.linalg.norm(xy[j] - xy[i])
    mut = xy + F / np.linalg.norm(F, axis=1)[:, None]
    assert np.abs(mut - d["xy_strict"][0]).max() > 1e-3
    xo, _, _ = oracle.minibatch(rp, ci, d["init"], 1, 256)
    assert np.array_equal(xo, d["xy_strict"][0])


def test_exact_repulsion_brute_force(oracle):
    pts = np.random.default_rng(9).random((500, 2))
    tg = np.array([0, 7, 499], np.int32)
    F = oracle.exact_repulsion(pts, tg)
    for a, i in enumerate(tg):
        d = pts - pts[i]
        r2 = (d ** 2).sum(1)
        r2[i] = np.inf
        assert np.allclose(F[a], (d / r2[:, None]).sum(0), rtol=1e-12, atol=0)


def test_energy_noise_band_fixture_is_consistent():
    # the final-energy band used by the GPU tests: spread of 1-ulp-perturbed FP64 runs
    for fx, lo in (("band_a2_grid64_b256_600.npz", 0.005), ("band_a8_grid64_b256_600.npz", 0.005),
                   ("band_a4_grid64_b256_600.npz", 0.05)):
        d = golden(fx)
        e = d["E"]
        spread = (e.max() - e.min()) / e[0]
        assert lo < spread < 0.5
