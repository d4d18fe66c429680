"""bl_run_batched: many graphs in one persistent launch, each laid out exactly as bl_run would
(algo 2 or 8).  Every graph is checked against the FP64 oracle run on that graph alone."""
import numpy as np
import pytest

from batchlayoutcode_amd import graphs
from conftest import diameter

pytestmark = pytest.mark.gpu


def _mixed_graphs(rng, sizes):
    out = []
    for k, nv in enumerate(sizes):
        kind = k % 3
        if kind == 0:
            r = max(2, int(np.sqrt(nv)))
            n, e = graphs.grid(r, max(2, nv // r))
        elif kind == 1:
            n, e = nv, graphs._spanning_tree_plus(nv, rng)
        else:
            n, e = nv, graphs._connected_er(nv, rng)
        rp, ci = graphs.csr_from_edges(n, e)
        out.append((rp, ci))
    return out


@pytest.mark.parametrize("algo,batch", [(2, 128), (8, 16), (2, 1000)])
def test_batched_matches_oracle(bl, oracle, algo, batch):
    rng = np.random.default_rng(7)
    sizes = [5, 17, 64, 100, 333, 700, 1500, 2048, 2900, 64, 90, 129]
    gs = _mixed_graphs(rng, sizes)
    xys = [np.random.default_rng(100 + k).random((len(rp) - 1, 2)) for k, (rp, _) in enumerate(gs)]
    print(f"[case] batched algo {algo} batch {batch}", flush=True)
    out, en, fin, st = bl.run_batched(gs, xys, algo=algo, iters=3, batch=batch)
    assert np.all(fin > 0) and st.seconds > 0
    for k, ((rp, ci), x0) in enumerate(zip(gs, xys)):
        ref, e_ref, _ = oracle.minibatch(rp, ci, x0, 3, batch, law=algo)
        err = np.abs(out[k] - ref).max() / diameter(ref)
        assert err <= 1e-8, (k, len(rp) - 1, err)
        assert abs(en[k] - e_ref) <= 1e-8 * e_ref


@pytest.mark.parametrize("all_edgeless", [False, True])
def test_batched_degenerate_graphs(bl, oracle, all_edgeless):
    """Graphs with no edges, a two-vertex graph and a graph whose last rows are empty, laid out
    in one launch beside an ordinary one; and a launch in which no graph has an edge."""
    rng = np.random.default_rng(13)
    gs = [(np.zeros(51, np.int32), np.zeros(0, np.int32)), (np.zeros(4, np.int32), np.zeros(0, np.int32))]
    if not all_edgeless:
        rp, ci = graphs.csr_from_edges(*graphs.grid(8, 8))
        tail = np.concatenate([rp, np.full(5, rp[-1], rp.dtype)])  # five isolated vertices at the end
        gs += [(np.array([0, 1, 2], np.int32), np.array([1, 0], np.int32)), (tail, ci),
               _mixed_graphs(rng, [100])[0]]
    xys = [np.random.default_rng(300 + k).random((len(rp) - 1, 2)) for k, (rp, _) in enumerate(gs)]
    print(f"[case] batched degenerate all_edgeless={all_edgeless}", flush=True)
    out, en, fin, _ = bl.run_batched(gs, xys, algo=2, iters=3, batch=16)
    for k, ((rp, ci), x0) in enumerate(zip(gs, xys)):
        ref, e_ref, _ = oracle.minibatch(rp, ci, x0, 3, 16, law=2)
        err = np.abs(out[k] - ref).max() / diameter(ref)
        assert err <= 1e-8, (k, len(rp) - 1, err)
        assert abs(en[k] - e_ref) <= 1e-8 * e_ref


def test_batched_global_memory_path_and_fp32(bl, oracle):
    # graphs too large for the LDS copy read their positions from global memory
    rng = np.random.default_rng(9)
    gs = _mixed_graphs(rng, [8000, 40, 6000])
    xys = [np.random.default_rng(200 + k).random((len(rp) - 1, 2)) for k, (rp, _) in enumerate(gs)]
    print("[case] batched large graphs", flush=True)
    out, en, fin, st = bl.run_batched(gs, xys, algo=2, iters=1, batch=128)
    out32, en32, _, _ = bl.run_batched(gs, xys, algo=2, iters=1, batch=128, precision=bl.PREC_FP32)
    for k, ((rp, ci), x0) in enumerate(zip(gs, xys)):
        ref, e_ref, _ = oracle.minibatch(rp, ci, x0, 1, 128)
        assert np.abs(out[k] - ref).max() <= 1e-10 * diameter(ref)
        assert np.abs(out32[k] - ref).max() <= 1e-4 * diameter(ref)


def test_batched_rejects_bad_arguments(bl):
    rp, ci = graphs.csr_from_edges(*graphs.grid(3, 3))
    x = np.zeros((9, 2))
    for kw in (dict(batch=2000), dict(precision=bl.PREC_FP64), dict(algo=4)):
        with pytest.raises(bl.BLError) as ei:
            bl.run_batched([(rp, ci)], [x], iters=1, **{"batch": 16, **kw})
        assert ei.value.code == -1
    big = graphs.csr_from_edges(*graphs.grid(100, 100))
    with pytest.raises(bl.BLError):
        bl.run_batched([big], [np.zeros((10000, 2))], iters=1, batch=128)


def test_batched_shapes_give_the_same_bits(bl, monkeypatch):
    """Every shape of a graph's batch step (64..1024 batch slots served at once, the rest of the
    1024 threads splitting the j range) and the autotuner's choice: the same positions and
    energies bit for bit, since the 16 fixed j chunks are always combined by the same pairwise
    tree."""
    rng = np.random.default_rng(21)
    gs = _mixed_graphs(rng, [70, 300, 1000, 2500, 4000, 130, 900])
    xys = [np.random.default_rng(400 + k).random((len(rp) - 1, 2)) for k, (rp, _) in enumerate(gs)]
    runs = {}
    for batch in (128, 1000):
        for shape in ("64", "128", "256", "512", "1024", "tuned", "rule"):
            for k in ("BL_BATCHED_IPW", "BL_TUNE"):
                monkeypatch.delenv(k, raising=False)
            if shape == "tuned":
                monkeypatch.setenv("BL_TUNE", "1")
            elif shape == "rule":
                monkeypatch.setenv("BL_TUNE", "0")
            else:
                monkeypatch.setenv("BL_BATCHED_IPW", shape)
            print(f"[case] batched shape {shape} batch {batch}", flush=True)
            runs[shape] = bl.run_batched(gs, xys, algo=2, iters=2, batch=batch)
        ref_xy, ref_en = runs["rule"][0], runs["rule"][1]
        for shape, (xy, en, _, _) in runs.items():
            assert all(np.array_equal(a, b) for a, b in zip(xy, ref_xy)), (batch, shape)
            assert np.array_equal(en, ref_en), (batch, shape)
