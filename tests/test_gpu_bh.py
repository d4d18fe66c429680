"""GPU parity of the Barnes-Hut path (algo 4): Morton scales, sort permutation and
quadtree bit-exact against the reference's own dumps; forces bit-exact (FP64 mode) or
within FP32 tolerance; algo 4 layouts bit-exact (FP64) / teacher-forced (FP32)."""
import numpy as np
import pytest

from batchlayoutcode_amd import graphs
from conftest import diameter, golden, golden_files, golden_input
from energyutil import energy_bars, perturbed
from treeutil import canonical_from_child_array, canonical_from_reference

pytestmark = pytest.mark.gpu
BH = golden_files("bh_*.npz")


@pytest.mark.parametrize("fx", BH)
def test_morton_scales_bit_exact(bl, fx):
    d = golden(fx)
    print(f"[case] morton {fx}", flush=True)
    xs, ys, mm = bl.bh_morton(d["pts"])
    assert mm == d["maxmin"]
    assert np.array_equal(xs, d["scales_strict"][:, 0]) and np.array_equal(ys, d["scales_strict"][:, 1])


@pytest.mark.parametrize("fx", BH)
def test_permutation_and_tree_bit_exact(bl, fx):
    d = golden(fx)
    print(f"[case] tree {fx}", flush=True)
    t = bl.bh_tree(d["pts"])
    assert np.array_equal(t["perm"], d["perm"])
    got = canonical_from_child_array(t["child"], t["diam"], t["cx"], t["cy"], t["root"])
    assert got == canonical_from_reference(d)


@pytest.mark.parametrize("fx", BH)
def test_forces(bl, fx):
    d = golden(fx)
    print(f"[case] forces {fx}", flush=True)
    f64 = bl.bh_forces(d["pts"], float(d["theta"]), precision=bl.PREC_FP64)
    assert np.array_equal(f64, d["force_strict"])
    f32 = bl.bh_forces(d["pts"], float(d["theta"]), precision=bl.PREC_FP32)
    err = np.abs(f32 - d["force_strict"]).max()
    scale = np.abs(d["force_strict"]).max()
    print(f"{fx}: FP32 force max abs err {err:.2e} (max |F| {scale:.1f})")
    assert err <= 2e-5 * max(1.0, scale)
    # the default precision: FP64 differences and sums, refined reciprocal square roots
    fmx = bl.bh_forces(d["pts"], float(d["theta"]), precision=bl.PREC_MIXED)
    assert np.abs(fmx - d["force_strict"]).max() <= 1e-12 * max(1.0, scale)


def test_forces_at_query_points_match_oracle(bl, oracle):
    rng = np.random.default_rng(4)
    pts = rng.random((3000, 2))
    q = rng.random((777, 2)) * 1.4 - 0.2
    for theta in (0.4, 1.2):
        ref = oracle.bh_forces(pts, theta, q)
        got = bl.bh_forces(pts, theta, q, precision=bl.PREC_FP64)
        assert np.array_equal(got, ref)


def test_duplicate_points_are_an_error(bl):
    pts = np.random.default_rng(2).random((100, 2))
    pts[17] = pts[3]
    with pytest.raises(bl.BLError) as ei:
        bl.bh_tree(pts)
    assert ei.value.code == -5


@pytest.mark.parametrize("n", [1023, 1025, 32769, 32770, 2097152, 2097153])
def test_sort_tile_and_scan_boundaries(bl, oracle, n):
    # the radix sort's tiles hold 1,024 keys (n = 1023, 1025: a partial tile); the node-id
    # scan over n - 1 boundary flags takes one workgroup up to 32,768 flags and three
    # launches beyond (n = 32769, 32770); the per-digit row scan works in chunks of 2,048
    # tiles (n = 2097152 fills one exactly, 2097153 starts a second): permutation and FP64
    # forces stay bit-exact
    pts = np.random.default_rng(n).random((n, 2))
    print(f"[case] sort boundaries n {n}", flush=True)
    t = bl.bh_tree(pts)
    ref = oracle.bh_tree(pts)
    assert np.array_equal(t["perm"], ref["perm"])
    q = pts[np.random.default_rng(1).choice(n, 64, replace=False)]
    got = bl.bh_forces(pts, 1.2, queries=q, precision=bl.PREC_FP64)
    assert np.array_equal(got, oracle.bh_forces(pts, 1.2, queries=q))


def test_sort_runs_equal_in_the_top_48_key_bits(bl, oracle):
    """The radix passes order the Morton keys by their top 48 bits (24 per axis); keys equal
    there, points within 2^-24 of the frame's side on both axes, are put in full-key order
    afterwards (k_sort_ties).  Pairs and triples of points 2^-28 apart in random directions,
    the points in shuffled order so that a run's ids are not already in key order: the
    permutation matches the oracle's full sort, and the FP64 forces stay bit-exact."""
    rng = np.random.default_rng(11)
    base = rng.random((20000, 2))

    def off(k):  # at least 2^-30 per axis: the 32-bit scales differ (distinct keys)
        return rng.choice([-1.0, 1.0], (k, 2)) * rng.uniform(0.25, 1.0, (k, 2)) * 2.0**-28

    pts = np.concatenate([base, base[:3000] + off(3000), base[:500] + off(500)])
    pts = pts[rng.permutation(len(pts))]
    xs, ys, _ = oracle.morton(pts)
    top = (ys.astype(np.uint64) >> np.uint64(8)) << np.uint64(24) | (xs.astype(np.uint64) >> np.uint64(8))
    _, cnt = np.unique(top, return_counts=True)
    full = ys.astype(np.uint64) << np.uint64(32) | xs.astype(np.uint64)
    assert len(np.unique(full)) == len(pts)  # no duplicate points
    print(f"[case] sort ties: {int((cnt > 1).sum())} runs, longest {int(cnt.max())}", flush=True)
    assert (cnt > 1).sum() >= 2000 and cnt.max() >= 3
    t = bl.bh_tree(pts)
    ref = oracle.bh_tree(pts)
    assert np.array_equal(t["perm"], ref["perm"])
    q = pts[rng.choice(len(pts), 64, replace=False)]
    got = bl.bh_forces(pts, 1.2, queries=q, precision=bl.PREC_FP64)
    assert np.array_equal(got, oracle.bh_forces(pts, 1.2, queries=q))


@pytest.mark.parametrize("n", [1, 2, 3, 5])
def test_tiny_point_sets(bl, oracle, n):
    pts = np.random.default_rng(n).random((n, 2))
    ref = oracle.bh_forces(pts, 1.2)
    got = bl.bh_forces(pts, 1.2, precision=bl.PREC_FP64)
    assert np.array_equal(got, ref)


A4 = golden_files("traj_a4_*.npz")


@pytest.mark.parametrize("fx", A4)
def test_algo4_fp64_matches_strict_reference(bl, fx):
    d = golden(fx)
    rp, ci = bl.mtx_read(golden_input(str(d["input"])))
    with bl.Graph(rp, ci) as g:
        for kk, k in enumerate(d["ks"]):
            print(f"[case] {fx} k={k} fp64", flush=True)
            xy, st = g.run(d["init"], algo=4, iters=int(k), batch=int(d["batch"]), bht=float(d["bht"]),
                           precision=bl.PREC_FP64)
            assert np.array_equal(xy, d["xy_strict"][kk]), (fx, k, np.abs(xy - d["xy_strict"][kk]).max())
            assert st.energy == d["E_strict"][kk]


@pytest.mark.parametrize("fx", A4)
def test_algo4_fp32_first_and_teacher_forced(bl, fx):
    d = golden(fx)
    rp, ci = bl.mtx_read(golden_input(str(d["input"])))
    ks = list(d["ks"])
    kw = dict(algo=4, iters=1, batch=int(d["batch"]), bht=float(d["bht"]), precision=bl.PREC_FP32)
    with bl.Graph(rp, ci) as g:
        xy, st = g.run(d["init"], **kw)
        err = np.abs(xy - d["xy_strict"][0]).max() / diameter(d["xy_strict"][0])
        print(f"{fx}: iteration 1 max|dx|/diam = {err:.3e}")
        assert err <= 1e-4
        for kk in range(1, len(ks)):
            if ks[kk] != ks[kk - 1] + 1:
                continue
            xy, st = g.run(d["xy_strict"][kk - 1], step0=0.999 ** ks[kk - 1], **kw)
            ref = d["xy_strict"][kk]
            err = np.abs(xy - ref).max() / diameter(ref)
            print(f"{fx}: {ks[kk-1]}->{ks[kk]} max|dx|/diam = {err:.3e}")
            assert err <= 1e-4


@pytest.mark.parametrize("case", ["edgeless", "n2", "n3_path"])
def test_algo4_degenerate_graphs_match_oracle(bl, oracle, case):
    """Barnes-Hut layouts of graphs with no edges (repulsion only) and of the smallest graphs:
    the FP64 path bit-exact with the oracle over 3 iterations, the default path within 1e-4 of
    the layout diameter after one."""
    rng = np.random.default_rng(12)
    if case == "edgeless":
        n = 300
        rp, ci = np.zeros(n + 1, np.int32), np.zeros(0, np.int32)
    elif case == "n2":
        n = 2
        rp, ci = np.array([0, 1, 2], np.int32), np.array([1, 0], np.int32)
    else:
        n = 3
        rp, ci = np.array([0, 1, 3, 4], np.int32), np.array([1, 0, 2, 1], np.int32)
    init = rng.random((n, 2))
    ref, e_ref = oracle.bh_layout(rp, ci, init, 3, 16, theta=1.2)
    ref1, _ = oracle.bh_layout(rp, ci, init, 1, 16, theta=1.2)
    print(f"[case] algo 4 degenerate {case}", flush=True)
    with bl.Graph(rp, ci) as g:
        xy, st = g.run(init, algo=4, iters=3, batch=16, bht=1.2, precision=bl.PREC_FP64)
        assert np.array_equal(xy, ref) and st.energy == e_ref
        xy1, _ = g.run(init, algo=4, iters=1, batch=16, bht=1.2)
    assert np.abs(xy1 - ref1).max() <= 1e-4 * diameter(ref1)


def test_algo4_final_energy_600_iterations(bl):
    """600 iterations of algo 4 on the 64x64 grid from the fixture init and 95 1-ulp
    perturbations of it, on the default and the FP64 path (energyutil.energy_bars)."""
    band = golden("band_a4_grid64_b256_600.npz")
    init = golden("traj_a4_grid64_b256.npz")["init"]
    rp, ci = bl.mtx_read(golden_input("grid64x64.mtx"))
    starts = 96
    e_def, e64 = [], []
    with bl.Graph(rp, ci) as g:
        print(f"[case] grid64 algo 4 600 iterations, {starts} starts, default and FP64", flush=True)
        for p in range(starts):
            xy0 = perturbed(init, p)
            e_def.append(g.run(xy0, algo=4, iters=600, batch=256)[1].energy)
            e64.append(g.run(xy0, algo=4, iters=600, batch=256, precision=bl.PREC_FP64)[1].energy)
    e_def, e64 = np.array(e_def), np.array(e64)
    dm, se_d, rb, se_b = energy_bars(e_def, e64, band["E"])
    print(f"algo 4: default - FP64 {dm:+.4f} (SE {se_d:.4f}); FP64 mean vs reference band {rb:+.4f} (SE {se_b:.4f}); "
          f"FP64 runs sd {e64.std(ddof=1) / e64.mean():.4f}")
    assert abs(dm) <= max(0.01, 3 * se_d)
    assert abs(rb) <= max(0.01, 3 * se_b)


def _hubby_graph(n, hubs, seed):
    """Random graph with a few hub vertices of very high degree (rows spanning many threads and
    workgroups of the batch walk)."""
    rng = np.random.default_rng(seed)
    e = [rng.integers(0, n, size=(3 * n, 2))]
    for h in rng.choice(n, hubs, replace=False):
        others = rng.choice(n, n // 3, replace=False)
        e.append(np.stack([np.full(len(others), h), others], 1))
    e = np.concatenate(e)
    e = e[e[:, 0] != e[:, 1]]
    return graphs.csr_from_edges(n, e)


@pytest.mark.parametrize("kind,algo,weighted,batch", [("grid", 4, False, 2048), ("hubs", 4, False, 2048),
                                                      ("hubs", 40, False, 2048), ("hubs", 4, True, 1000),
                                                      ("dense", 4, True, 2048), ("dense", 40, False, 1500)])
def test_multi_workgroup_walk_matches_single(bl, monkeypatch, kind, algo, weighted, batch):
    # batches of 1000..2048 vertices carry 8k..130k CSR entries: the walk spreads them over
    # several workgroups (on "dense" a thread sums more entries than it loads a batch ahead);
    # it must agree with the one-workgroup walk up to the order of partial sums
    if kind == "grid":
        rp, ci = graphs.csr_from_edges(*graphs.grid(100, 100))
    elif kind == "hubs":
        rp, ci = _hubby_graph(12000, 4, 3)
    else:
        rng = np.random.default_rng(5)
        e = rng.integers(0, 12000, size=(12000 * 30, 2))
        rp, ci = graphs.csr_from_edges(12000, e[e[:, 0] != e[:, 1]])
    n = len(rp) - 1
    w = np.random.default_rng(4).uniform(0.5, 2.0, len(ci)).astype(np.float32) if weighted else None
    x0 = np.random.default_rng(2).random((n, 2))
    print(f"[case] grid walk {kind} algo {algo} weighted {weighted} batch {batch} n {n} nnz {len(ci)}", flush=True)
    # iteration 1 from x0, then iteration 2 teacher-forced from the one-workgroup walk's
    # iteration-1 layout: free-running, the order of the partial sums alone can grow to 1e-8
    # by iteration 2 on the dense graph (a vertex whose force nearly cancels)
    start = x0
    for it in range(2):
        out = {}
        for mode in ("single", "grid"):
            monkeypatch.setenv("BL_WALK", mode)
            with bl.Graph(rp, ci, w) as g:
                out[mode], st = g.run(start, algo=algo, iters=1, batch=batch, bht=1.2, step0=0.999 ** it)
                out[mode + "_E"] = st.energy
        d = diameter(out["single"])
        assert np.abs(out["grid"] - out["single"]).max() <= 1e-10 * d, it
        assert abs(out["grid_E"] - out["single_E"]) <= 1e-10 * out["single_E"], it
        start = out["single"]


def _lag_steps(rp, ci, B):
    """The lagged walk's steps for batches of B: consecutive batches grouped while none of a
    joining batch's vertices has a neighbour in the group (at most 512 // B batches), when that
    makes at most four fifths as many steps as batches; else one batch a step."""
    n = len(rp) - 1
    nb = (n + B - 1) // B
    rows = np.repeat(np.arange(n), np.diff(rp))
    bi, bj = rows // B, ci // B
    early = bj < bi
    last_early = np.full(nb, -1)  # the largest neighbour id of a batch's vertices below the batch
    np.maximum.at(last_early, bi[early], ci[early])

    def steps(T):
        gmax, ng, cur0, size = max(1, T // B), 1, 0, 1
        for b in range(1, nb):
            if size < gmax and last_early[b] < cur0:
                size += 1
            else:
                ng, cur0, size = ng + 1, b * B, 1
        return ng

    g512 = steps(512)
    return g512 if 5 * g512 <= 4 * nb else steps(256)


@pytest.mark.parametrize("kind,algo,weighted,batch", [("hubs", 4, False, 1024), ("hubs", 40, False, 1000),
                                                      ("hubs", 4, True, 256), ("rmat", 4, True, 1024),
                                                      ("rmat", 4, False, 512), ("dense", 4, True, 1024)])
def test_wide_walk_matches_single(bl, monkeypatch, kind, algo, weighted, batch):
    """The wide lagged walk (a walking workgroup sums the neighbours moved in the last D_b - 1
    batches from an LDS ring, helper workgroups the others in chunks of rows) agrees with the
    one-workgroup walk up to the order of partial sums, at iteration 1 and at iteration 2
    teacher-forced from the one-workgroup walk's iteration 1; its layout is the same bits
    whatever the number of helpers (one helper makes the walker wait at almost every batch),
    and a smaller largest lag moves entries between the walker and the helpers without changing
    the result beyond the order of sums.  R-MAT and the last partial batches (21024 vertices)
    cover hubs in the first batches and the threads past a batch's end."""
    if kind == "hubs":
        rp, ci = _hubby_graph(12000, 4, 3)
    elif kind == "rmat":
        n0, e0, _ = graphs.rmat(15, 8, seed=2)
        rp, ci = graphs.csr_from_edges(n0, e0)
    else:
        rng = np.random.default_rng(5)
        e = rng.integers(0, 12000, size=(12000 * 30, 2))
        rp, ci = graphs.csr_from_edges(12000, e[e[:, 0] != e[:, 1]])
    n = len(rp) - 1
    w = np.random.default_rng(4).uniform(0.5, 2.0, len(ci)).astype(np.float32) if weighted else None
    x0 = np.random.default_rng(2).random((n, 2))
    print(f"[case] wide walk {kind} algo {algo} weighted {weighted} batch {batch} n {n} nnz {len(ci)}", flush=True)
    runs = (("single", dict(BL_WALK="single")), ("wide", dict(BL_WALK="wide")),
            ("wide_h1", dict(BL_WALK="wide", BL_WIDE_H="1")), ("wide_h3", dict(BL_WALK="wide", BL_WIDE_H="3")),
            ("wide_d5", dict(BL_WALK="wide", BL_WIDE_D="5")))
    start = x0
    for it in range(2):
        out = {}
        for mode, env in runs:
            for k in ("BL_WALK", "BL_WIDE_H", "BL_WIDE_D"):
                monkeypatch.delenv(k, raising=False)
            for k, v in env.items():
                monkeypatch.setenv(k, v)
            with bl.Graph(rp, ci, w) as g:
                out[mode], st = g.run(start, algo=algo, iters=1, batch=batch, bht=1.2, step0=0.999 ** it)
            out[mode + "_E"] = st.energy
            if mode != "single":
                assert st.walk == bl.WALK_WIDE, (mode, st.walk)
                assert st.walk_lag == (5 if mode == "wide_d5" else 16), (mode, st.walk_lag)
                print(f"  it {it} {mode}: batches ready when reached {st.walk_helped} of {(n + batch - 1) // batch}")
        d = diameter(out["single"])
        for mode in ("wide", "wide_d5"):
            err = np.abs(out[mode] - out["single"]).max() / d
            print(f"  it {it} {mode}: {err:.2e} of the diameter")
            assert err <= 1e-10, (it, mode)
            assert abs(out[mode + "_E"] - out["single_E"]) <= 1e-10 * out["single_E"], (it, mode)
        for mode in ("wide_h1", "wide_h3"):
            assert np.array_equal(out[mode], out["wide"]) and out[mode + "_E"] == out["wide_E"], (it, mode)
        start = out["single"]


@pytest.mark.parametrize("kind,weighted,batch,algo,prec", [("rgg", False, 256, 4, 0), ("rgg", True, 256, 4, 0),
                                                           ("rgg", False, 100, 4, 0), ("rgg", True, 37, 4, 0),
                                                           ("rgg", False, 256, 40, 0), ("rgg", False, 256, 4, 2),
                                                           ("grid", False, 256, 4, 0), ("small", True, 256, 4, 0),
                                                           ("small", False, 1, 4, 0), ("isolated", False, 64, 4, 0)])
def test_row_walk_matches_entry_split_walk(bl, monkeypatch, kind, weighted, batch, algo, prec):
    # short rows and batches of at most 256 vertices take the one-thread-per-vertex walk (the
    # attraction of neighbours in the same or later batches summed beforehand); it must agree
    # with the entry-split walk up to the order of partial sums
    if kind == "grid":
        rp, ci = graphs.csr_from_edges(*graphs.grid(64, 64))
    elif kind == "small":  # one partial batch at B=256, or a batch per vertex
        rp, ci = graphs.csr_from_edges(*graphs.rgg(120, seed=6))
    elif kind == "isolated":  # a random geometric graph with a sparse radius: many empty rows
        n0 = 3000
        pts = np.random.default_rng(9).random((n0, 2))
        d2 = ((pts[:, None, :] - pts[None, :, :]) ** 2).sum(-1)
        a, b = np.nonzero(np.triu(d2 < 0.012 ** 2, 1))
        rp, ci = graphs.csr_from_edges(n0, np.stack([a, b], 1))
        assert (np.diff(rp) == 0).sum() > 100
    else:
        rp, ci = graphs.csr_from_edges(*graphs.rgg(20000, seed=5))
    n = len(rp) - 1
    w = np.random.default_rng(7).uniform(0.5, 2.0, len(ci)).astype(np.float32) if weighted else None
    x0 = np.random.default_rng(8).random((n, 2))
    print(f"[case] row walk {kind} weighted {weighted} batch {batch} algo {algo} precision {prec}", flush=True)
    out = {}
    for mode in ("single", "pf"):
        monkeypatch.setenv("BL_WALK", mode)
        with bl.Graph(rp, ci, w) as g:
            out[mode], st = g.run(x0, algo=algo, iters=3, batch=batch, bht=1.2, precision=prec)
            out[mode + "_E"] = st.energy
    d = diameter(out["pf"])
    assert np.abs(out["single"] - out["pf"]).max() <= 1e-10 * d
    assert abs(out["single_E"] - out["pf_E"]) <= 1e-10 * out["pf_E"]


@pytest.mark.parametrize("kind,weighted,batch,algo", [("rgg", False, 256, 4), ("rgg", True, 256, 4), ("rgg", False, 100, 4),
                                                      ("grid", False, 256, 4), ("delaunay", False, 256, 40),
                                                      ("delaunay", True, 200, 4), ("delaunay", False, 64, 4),
                                                      ("delaunay", True, 100, 40)])
def test_lag_walk_matches_row_walk_and_is_deterministic(bl, monkeypatch, kind, weighted, batch, algo):
    """The lagged walk (helper workgroups sum the neighbours moved at least D batches earlier; the
    walker sums the rest and forms a batch's sum itself when its helper is late) against the
    one-workgroup row walk, and bit for bit against itself with no helpers at all and with other
    lags and publishing periods: whichever workgroup formed a batch's sum, it is the same.
    Against the row walk the partial sums come in another order and |d| and F/|F| come from a
    refined reciprocal square root instead of an IEEE square root and division, so the two agree
    to rounding: 1e-10 of the diameter after the first iteration, and after a third iteration
    teacher-forced from the row walk's second (free-running, such differences grow chaotically:
    2e-9 after three iterations).  At batches of 64 and 100 on the Delaunay graph most
    consecutive batches share no edge, and the walk moves groups of up to 512 vertices a step."""
    if kind == "grid":  # natural order: every early neighbour is in the previous batch
        rp, ci = graphs.csr_from_edges(*graphs.grid(160, 160))
    elif kind == "delaunay":
        rp, ci = graphs.csr_from_edges(*graphs.delaunay_holes(200_000, seed=3))
    else:  # at most 4 neighbours of a vertex in the 19 batches before its own (lag 20)
        rp, ci = graphs.csr_from_edges(*graphs.rgg(250_000, seed=5))
    n = len(rp) - 1
    w = np.random.default_rng(7).uniform(0.5, 2.0, len(ci)).astype(np.float32) if weighted else None
    x0 = np.random.default_rng(8).random((n, 2))
    print(f"[case] lag walk {kind} weighted {weighted} batch {batch} algo {algo} n {n}", flush=True)
    out, helped = {}, {}
    # a short lag has the helpers read positions the walker has just published
    for mode, env in (("row", dict(BL_WALK="row")), ("lag", {}), ("lag_h0", dict(BL_LAG_H="0")),
                      ("lag_h4", dict(BL_LAG_H="4")), ("lag_d12", dict(BL_LAG_D="12")),
                      ("lag_d12_h0", dict(BL_LAG_D="12", BL_LAG_H="0"))):
        for k in ("BL_WALK", "BL_LAG_H", "BL_LAG_D"):
            monkeypatch.delenv(k, raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        with bl.Graph(rp, ci, w) as g:
            out[mode], st = g.run(x0, algo=algo, iters=3, batch=batch, bht=1.2)
            out[mode + "_E"] = st.energy
            helped[mode] = st.walk_helped
            assert st.walk == (bl.WALK_ROW if mode == "row" else bl.WALK_LAG), (mode, st.walk)
            if mode in ("row", "lag", "lag_d12"):  # iteration 1, and iteration 3 teacher-forced
                out[mode + "_1"], _ = g.run(x0, algo=algo, iters=1, batch=batch, bht=1.2)
                if mode == "row":
                    x2, _ = g.run(x0, algo=algo, iters=2, batch=batch, bht=1.2)
                out[mode + "_3tf"], _ = g.run(x2, algo=algo, iters=1, batch=batch, bht=1.2, step0=0.999 ** 2)
    nb = (n + batch - 1) // batch
    ng = _lag_steps(rp, ci, batch)
    print(f"batches {nb}, walk steps (groups) {ng}, formed by the helpers (last iteration): {helped}")
    if batch <= 100 and kind == "delaunay":
        assert ng < nb // 2  # groups of several batches
    d = diameter(out["row"])
    for m in ("lag", "lag_d12"):
        e1 = np.abs(out[m + "_1"] - out["row_1"]).max() / d
        e3 = np.abs(out[m + "_3tf"] - out["row_3tf"]).max() / d
        print(f"{m} vs row: iteration 1 {e1:.2e}, teacher-forced iteration 3 {e3:.2e}, free-running 3 iterations "
              f"{np.abs(out[m] - out['row']).max() / d:.2e} of the diameter")
        assert e1 <= 1e-10 and e3 <= 1e-10, m
    assert abs(out["lag_E"] - out["row_E"]) <= 1e-6 * out["row_E"]
    assert helped["lag_h0"] == 0 and helped["lag_d12_h0"] == 0
    assert helped["lag"] > ng // 2 and helped["lag_d12"] > 0
    assert np.array_equal(out["lag"], out["lag_h0"]) and out["lag_E"] == out["lag_h0_E"]
    assert np.array_equal(out["lag"], out["lag_h4"]) and out["lag_E"] == out["lag_h4_E"]
    assert np.array_equal(out["lag_d12"], out["lag_d12_h0"]) and out["lag_d12_E"] == out["lag_d12_h0_E"]


def test_lag_walk_declines_crowded_recent_batches(bl, monkeypatch):
    """A vertex with more than four neighbours in the batches just before its own (40k-vertex
    Delaunay graph, batches of 256, lag 20) does not fit the walk's stage: the row walk runs.  A
    shorter lag that fits takes the lagged walk again, and so does the default, which picks the
    largest lag that fits: every vertex's fifth-nearest earlier batch among its neighbours
    (computed here from the CSR) is at least that far back."""
    rp, ci = graphs.csr_from_edges(*graphs.delaunay_holes(40_000, seed=3))
    n, B = len(rp) - 1, 256
    x0 = np.random.default_rng(8).random((n, 2))
    # the largest lag with at most four neighbours in the lag - 1 batches before a vertex's own
    rows = np.repeat(np.arange(n), np.diff(rp))
    gaps = rows // B - ci // B
    order = np.lexsort((gaps, rows))
    r_s, g_s = rows[order], gaps[order]
    keep = g_s >= 1
    r_s, g_s = r_s[keep], g_s[keep]
    first = np.searchsorted(r_s, np.arange(n))
    cnt = np.bincount(r_s, minlength=n)
    assert _lag_steps(rp, ci, B) == (n + B - 1) // B  # every consecutive pair shares an edge: no groups
    d_fit = min(48, ((n + B - 1) // B) // 2, int(g_s[first[cnt >= 5] + 4].min()) if (cnt >= 5).any() else 48)
    for k in ("BL_WALK", "BL_LAG_H", "BL_LAG_D"):
        monkeypatch.delenv(k, raising=False)
    outs = {}
    for lag, walk in (("20", bl.WALK_ROW), ("6", bl.WALK_LAG), (None, bl.WALK_LAG)):
        if lag:
            monkeypatch.setenv("BL_LAG_D", lag)
        else:
            monkeypatch.delenv("BL_LAG_D", raising=False)
        with bl.Graph(rp, ci) as g:
            outs[lag], st = g.run(x0, algo=4, iters=1, batch=B, bht=1.2)
        print(f"lag {lag}: walk {st.walk}, lag taken {st.walk_lag} (largest that fits: {d_fit})")
        assert st.walk == walk, lag
        assert st.walk_lag == (0 if walk == bl.WALK_ROW else int(lag) if lag else d_fit), lag
    assert d_fit < 20
    for lag in ("6", None):
        assert np.abs(outs[lag] - outs["20"]).max() <= 1e-10 * diameter(outs["20"])


@pytest.mark.parametrize("algo", [2, 4, 40])
def test_nonfinite_guard(bl, oracle, algo):
    """A vertex whose force is exactly zero (a graph of one vertex, or a vertex placed where every
    force cancels) moves to NaN, as the reference's x += STEP F/|F| does (Coordinate.h:63-65); the
    run keeps that result and reports the vertices whose final position is not finite."""
    one = (np.array([0, 0], np.int32), np.zeros(0, np.int32))
    print(f"[case] non-finite guard algo {algo}", flush=True)
    with bl.Graph(*one) as g:
        xy, st = g.run(np.array([[0.3, 0.7]]), algo=algo, iters=2, batch=16, bht=1.2)
    assert np.isnan(xy).all() and st.nonfinite == 1
    if algo == 2:
        ref, _, _ = oracle.minibatch(*one, np.array([[0.3, 0.7]]), 2, 16, law=2)
        assert np.isnan(ref).all()
    rp, ci = graphs.csr_from_edges(*graphs.grid(8, 8))
    with bl.Graph(rp, ci) as g:
        xy, st = g.run(np.random.default_rng(1).random((64, 2)), algo=algo, iters=3, batch=16, bht=1.2)
    assert np.isfinite(xy).all() and st.nonfinite == 0
    out, en, _, stb = bl.run_batched([one, (rp, ci)], [np.array([[0.3, 0.7]]), np.random.default_rng(1).random((64, 2))],
                                      algo=2, iters=2, batch=16)
    assert np.isnan(out[0]).all() and np.isfinite(out[1]).all() and stb.nonfinite == 1
