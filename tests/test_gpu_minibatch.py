"""GPU parity of the minibatch family (algo 2, 3, 8, converged algo 2, algo 6) through the C ABI.

Bars:
  * FP64 mode: coordinates and energy equal the strict-IEEE reference build bit for bit
    (LinLog: within the spread between the reference's fast-math and strict builds, since
    device log2 and glibc log2 may differ by an ulp).
  * mixed mode (the default: FP64 pair arithmetic with an FP32-seeded reciprocal): iteration 1
    from the fixture init, and every iteration 2..10 started from the reference's own state
    (teacher forcing), within max|dx| <= 1e-4 * layout diameter; final energy after 600
    iterations within the spread of FP64 runs from perturbed starts.
  * FP32 mode (FP32 |d|^2 and reciprocal): iteration 1 within 1e-4 * layout diameter.
"""
import numpy as np
import pytest

from conftest import diameter, golden, golden_files, golden_input
from energyutil import energy_bars, perturbed

pytestmark = pytest.mark.gpu
TOL = 1e-4  # max |dx| / layout diameter


def graph_of(bl, name, w=None):
    rp, ci = bl.mtx_read(golden_input(name))
    return bl.Graph(rp, ci, w)


def fixture_kw(d, k, precision):
    algo = int(d["algo"])
    kw = dict(batch=int(d["batch"]), weight=float(d["weight"]), bht=float(d["bht"]), precision=precision)
    if algo == 21:
        kw.update(algo=2, converged=1, engt=float(d["engt"]), iters=0)
    else:
        kw.update(algo=algo, iters=int(k))
    return kw


MINIBATCH = [f for f in golden_files("traj_*.npz") if not f.startswith("traj_a4_")]


@pytest.mark.parametrize("fx", MINIBATCH)
def test_fp64_mode_matches_strict_reference(bl, fx):
    d = golden(fx)
    with graph_of(bl, str(d["input"])) as g:
        for kk, k in enumerate(d["ks"]):
            print(f"[case] {fx} k={k} fp64", flush=True)
            xy, st = g.run(d["init"], **fixture_kw(d, k, bl.PREC_FP64))
            if int(d["algo"]) == 3:
                # device log2 and glibc log2 may differ by an ulp; that difference grows along the
                # trajectory like any FP64 perturbation, so the bar is the spread between the
                # reference's own fast-math and strict builds at the same iteration
                spread = np.abs(d["xy"][kk] - d["xy_strict"][kk]).max()
                assert np.abs(xy - d["xy_strict"][kk]).max() <= max(1e-12 * diameter(d["xy_strict"][kk]), spread)
                e_spread = abs(d["E"][kk] - d["E_strict"][kk])
                assert abs(st.energy - d["E_strict"][kk]) <= max(1e-12 * d["E_strict"][kk], e_spread)
            else:
                assert np.array_equal(xy, d["xy_strict"][kk]), (fx, k, np.abs(xy - d["xy_strict"][kk]).max())
                assert st.energy == d["E_strict"][kk]


@pytest.mark.parametrize("prec", ["mixed", "fp32"])
@pytest.mark.parametrize("fx", [f for f in MINIBATCH if "a21" not in f and "a6" not in f])
def test_first_iteration(bl, fx, prec):
    d = golden(fx)
    k = int(d["ks"][0])
    assert k == 1
    with graph_of(bl, str(d["input"])) as g:
        print(f"[case] {fx} k={k} {prec}", flush=True)
        xy, st = g.run(d["init"], **fixture_kw(d, k, bl.PREC_MIXED if prec == "mixed" else bl.PREC_FP32))
    ref = d["xy_strict"][0]
    err = np.abs(xy - ref).max() / diameter(ref)
    print(f"{fx}: iteration-{k} {prec} max|dx|/diam = {err:.3e}")
    assert err <= TOL
    assert abs(st.energy - d["E_strict"][0]) <= 1e-5 * d["E_strict"][0]


@pytest.mark.parametrize("fx,decay", [("traj_a2_grid64_b256.npz", 0.999), ("traj_a8_grid64_b256.npz", 0.99),
                                      ("traj_a2_grid8_b16.npz", 0.999), ("traj_a2_jagmesh1_b256.npz", 0.999)])
def test_mixed_teacher_forced_iterations(bl, fx, decay):
    d = golden(fx)
    ks = list(d["ks"])
    with graph_of(bl, str(d["input"])) as g:
        for kk in range(1, len(ks)):
            k0, k1 = ks[kk - 1], ks[kk]
            if k1 != k0 + 1:
                continue
            print(f"[case] {fx} teacher-forced {k0}->{k1}", flush=True)
            xy, st = g.run(d["xy_strict"][kk - 1], algo=int(d["algo"]), iters=1, batch=int(d["batch"]),
                           weight=float(d["weight"]), step0=decay ** k0, precision=bl.PREC_MIXED)
            ref = d["xy_strict"][kk]
            err = np.abs(xy - ref).max() / diameter(ref)
            rel_e = abs(st.energy - d["E_strict"][kk]) / d["E_strict"][kk]
            print(f"{fx}: {k0}->{k1} max|dx|/diam = {err:.3e}  energy rel = {rel_e:.2e}")
            assert err <= TOL
            assert rel_e <= 1e-4


@pytest.mark.parametrize("algo,fx", [(2, "band_a2_grid64_b256_600.npz"), (8, "band_a8_grid64_b256_600.npz")])
def test_final_energy_600_iterations(bl, algo, fx):
    """600 iterations on the 64x64 grid from the fixture init and 95 1-ulp perturbations of it,
    on the default and the FP64 path; bars at max(1 %, 3 standard errors) on the paired
    difference and on the FP64 mean against the reference's own runs (energyutil.energy_bars)."""
    band = golden(fx)
    init = golden("traj_a2_grid64_b256.npz")["init"]
    starts = 96
    e_def, e64 = [], []
    with graph_of(bl, "grid64x64.mtx") as g:
        print(f"[case] grid64 algo {algo} 600 iterations, {starts} starts, default and FP64", flush=True)
        for p in range(starts):
            xy0 = perturbed(init, p)
            xy, st = g.run(xy0, algo=algo, iters=600, batch=256)
            e_def.append(st.energy)
            assert st.iters_done == 600 and st.step_final == pytest.approx(
                0.999 ** 600 if algo == 2 else 0.99 ** 600, rel=1e-12)
            e64.append(g.run(xy0, algo=algo, iters=600, batch=256, precision=bl.PREC_FP64)[1].energy)
    e_def, e64 = np.array(e_def), np.array(e64)
    dm, se_d, rb, se_b = energy_bars(e_def, e64, band["E"])
    print(f"algo {algo}: default - FP64 {dm:+.5f} (SE {se_d:.5f}); FP64 mean vs reference band {rb:+.5f} "
          f"(SE {se_b:.5f}); FP64 runs sd {e64.std(ddof=1) / e64.mean():.4f}")
    assert abs(dm) <= max(0.01, 3 * se_d)
    assert abs(rb) <= max(0.01, 3 * se_b)


def test_fp32_run_is_deterministic(bl):
    d = golden("traj_a2_jagmesh1_b256.npz")
    with graph_of(bl, "jagmesh1.mtx") as g:
        a, ea = g.run(d["init"], algo=2, iters=5, batch=64)
        b, eb = g.run(d["init"], algo=2, iters=5, batch=64)
    assert np.array_equal(a, b) and ea.energy == eb.energy


@pytest.mark.parametrize("algo,prec,conv", [(2, 0, 0), (8, 2, 0), (3, 0, 0), (2, 1, 0), (2, 0, 1)])
def test_graph_replay_matches_launches(bl, monkeypatch, algo, prec, conv):
    """One iteration captured as a hipGraph and replayed (BL_GRAPH=1) against the same launches
    made one by one (the default): bit for bit, over runs of 1..5 replays made one after another with the
    host reading the results between them (the converged mode also reads the energy back after
    every replay), and with a run of the other kind in between."""
    rp, ci = bl.mtx_read(golden_input("jagmesh1.mtx"))
    x0 = np.random.default_rng(algo + 10 * prec).random((len(rp) - 1, 2))
    kw = dict(algo=algo, batch=64, precision=prec)
    if conv:
        kw.update(converged=1, engt=0.02)
    print(f"[case] graph replay algo {algo} precision {prec} converged {conv}", flush=True)
    with bl.Graph(rp, ci) as g:
        for iters in ((0,) if conv else (1, 2, 3, 5)):
            out = {}
            for mode in ("0", "1", "0"):
                monkeypatch.setenv("BL_GRAPH", mode)
                xy, st = g.run(x0, iters=iters, **kw)
                if mode in out:
                    assert np.array_equal(xy, out[mode][0]) and st.energy == out[mode][1].energy
                out[mode] = (xy, st)
            (a, sa), (b, sb) = out["0"], out["1"]
            assert sa.iters_done == sb.iters_done and sa.iters_done >= max(iters, 1)
            assert np.array_equal(a, b), (iters, np.abs(a - b).max())
            assert sa.energy == sb.energy and sa.step_final == sb.step_final


def _random_graph(rng, n, m, empty_rows=(), dup=False):
    e = rng.integers(0, n, size=(m, 2))
    e = e[e[:, 0] != e[:, 1]]
    rows = np.concatenate([e[:, 0], e[:, 1]])
    cols = np.concatenate([e[:, 1], e[:, 0]])
    keep = ~np.isin(rows, empty_rows) & ~np.isin(cols, empty_rows)
    rows, cols = rows[keep], cols[keep]
    if not dup:
        rc = np.unique(np.stack([rows, cols], 1), axis=0)
        rows, cols = rc[:, 0], rc[:, 1]
    o = np.lexsort((cols, rows))
    rows, cols = rows[o], cols[o]
    rp = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(rows, minlength=n), out=rp[1:])
    return rp.astype(np.int32), cols.astype(np.int32)


@pytest.mark.parametrize("case", ["isolated_rows", "duplicates", "weights", "b1", "b3", "b_gt_n", "selfloop",
                                  "edgeless", "n2"])
def test_edge_cases_match_oracle(bl, oracle, case):
    rng = np.random.default_rng(11)
    n, B, w = 70, 16, None
    if case == "edgeless":  # nnz == 0: repulsion only, every row empty
        rp, ci = np.zeros(n + 1, np.int32), np.zeros(0, np.int32)
    elif case == "n2":  # the smallest graph with a force: one edge, a batch wider than the graph
        n = 2
        rp, ci = np.array([0, 1, 2], np.int32), np.array([1, 0], np.int32)
    elif case == "isolated_rows":
        rp, ci = _random_graph(rng, n, 150, empty_rows=(5, 6, 40, n - 1))
    elif case == "duplicates":
        rp, ci = _random_graph(rng, n, 150, dup=True)
        rp2, ci2 = bl.mtx_read(golden_input("sym_duplicate.mtx"))
        assert len(ci2) == 10
    elif case == "selfloop":
        rp, ci = bl.mtx_read(golden_input("general_selfloop.mtx"))
        n = len(rp) - 1
        B = 2
    else:
        rp, ci = _random_graph(rng, n, 150)
        if case == "weights":
            w = rng.uniform(0.5, 2.0, size=len(ci)).astype(np.float32)
        B = {"b1": 1, "b3": 3, "b_gt_n": 1000}.get(case, B)
    init = rng.random((n, 2))
    for law in (2, 8):
        ref, e_ref, _ = oracle.minibatch(rp, ci, init, 4, B, law=law, weight=1.5, w_edge=w)
        with bl.Graph(rp, ci, w) as g:
            print(f"[case] edge {case} law {law}", flush=True)
            xy, st = g.run(init, algo=law, iters=4, batch=B, weight=1.5, precision=bl.PREC_FP64)
            assert np.array_equal(xy, ref) and st.energy == e_ref, (case, law)
            xy32, st32 = g.run(init, algo=law, iters=1, batch=B, weight=1.5, precision=bl.PREC_MIXED)
        ref1, _, _ = oracle.minibatch(rp, ci, init, 1, B, law=law, weight=1.5, w_edge=w)
        assert np.abs(xy32 - ref1).max() <= TOL * diameter(ref1)


def test_rgg_250k_one_iteration_fp32_vs_fp64(bl):
    """Full-size random geometric graph (250,000 points, seed 1): one iteration of the mixed
    and FP32 paths against the reference-order FP64 path on the same GPU."""
    from batchlayoutcode_amd import graphs

    n, e = graphs.rgg(250_000, seed=1)
    rp, ci = graphs.csr_from_edges(n, e)
    init = graphs.uniform_init(n, 0)
    with bl.Graph(rp, ci) as g:
        print("[case] rgg250k fp64 1 iteration", flush=True)
        x64, s64 = g.run(init, algo=2, iters=1, batch=256, precision=bl.PREC_FP64)
        print("[case] rgg250k mixed / fp32 1 iteration", flush=True)
        xm, sm = g.run(init, algo=2, iters=1, batch=256, precision=bl.PREC_MIXED)
        x32, s32 = g.run(init, algo=2, iters=1, batch=256, precision=bl.PREC_FP32)
    for tag, x, st in (("mixed", xm, sm), ("fp32", x32, s32)):
        err = np.abs(x - x64).max() / diameter(x64)
        rel_e = abs(st.energy - s64.energy) / s64.energy
        print(f"rgg250k {tag}: n={n} max|dx|/diam={err:.3e} energy rel={rel_e:.2e} "
              f"{st.seconds*1e3:.1f} ms (fp64 {s64.seconds*1e3:.1f} ms)")
        assert err <= TOL and rel_e <= 1e-5


@pytest.mark.parametrize("law,prec", [(2, 0), (8, 0), (2, 2), (8, 2)])
def test_hub_corrections_in_pieces(bl, oracle, law, prec):
    """Vertices with thousands of neighbours: their corrections are split over several waves
    (pieces of 256 entries) and summed per vertex; one iteration against the oracle."""
    rng = np.random.default_rng(21)
    n = 3000
    e = [rng.integers(0, n, size=(4 * n, 2))]
    for h in (0, 7, 1500, n - 1):  # hubs of 300..2000 neighbours, in different batches
        others = rng.choice(n, int(rng.integers(300, 2000)), replace=False)
        e.append(np.stack([np.full(len(others), h), others], 1))
    e = np.concatenate(e)
    e = e[e[:, 0] != e[:, 1]]
    rows = np.concatenate([e[:, 0], e[:, 1]])
    cols = np.concatenate([e[:, 1], e[:, 0]])
    rc = np.unique(np.stack([rows, cols], 1), axis=0)
    rp = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(rc[:, 0], minlength=n), out=rp[1:])
    rp, ci = rp.astype(np.int32), rc[:, 1].astype(np.int32)
    assert np.diff(rp).max() > 1000
    w = rng.uniform(0.5, 2.0, size=len(ci)).astype(np.float32)
    init = rng.random((n, 2))
    ref, _, _ = oracle.minibatch(rp, ci, init, 1, 256, law=law, weight=1.0, w_edge=w)
    with bl.Graph(rp, ci, w) as g:
        print(f"[case] hub corrections law {law} precision {prec}", flush=True)
        xy, st = g.run(init, algo=law, iters=1, batch=256, precision=prec)
    assert np.abs(xy - ref).max() <= TOL * diameter(ref)
