"""Named layout workloads (the benchmark configurations) for timing and profiling runs.

    python -m batchlayoutcode_amd.workloads <name> [--iters K] [--profile] [--precision P] [--calib-copy]

  grid64-algo2     64x64 grid (4,096 vertices), algo 2, -batch 256
  rgg250k-algo2    random geometric graph, 250k points (seed 1), mean degree 8, LCC; algo 2, -batch 256
  rgg250k-algo4    the same graph, Barnes-Hut (algo 4), -batch 256, -bht 1.2
  rmat20-algo4     R-MAT scale 20, edge factor 8, (0.57, 0.19, 0.19, 0.05), seed 2, LCC, FP32 weights
                   in [0.5, 2]; algo 4, -batch 1024, -bht 1.2
  rmat20-algo8     the same graph, algo 8, -batch 1024
  delaunay2m-far   Delaunay edges of 2M points minus 20 disks (seed 3), LCC; physically consistent
                   Barnes-Hut: monopole, quadrupole and learned far field (model trained first), -bht 1.0
  delaunay2m-algo40  the same graph, monopole Barnes-Hut alone at -bht 0.4 (no model: a profile of the
                   plain path with nothing else in the process)
  batched4096      4,096 small graphs (64..8192 vertices, seed 4), algo 2, -batch 128, one launch
  batched4096-pergraph  the same graphs laid out one after another, each by the single-graph path
                   (bl_run_device on device-resident positions): the baseline of the one launch

Initial coordinates are uniform in [0,1)^2 (seed 0).  Prints one JSON line per run.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import api, graphs

# the library's kernel classes (bl_stats.kernel_ms)
CLASSES = ("pairs_or_traversal", "move_or_walk", "energy", "tree", "cell_summaries", "walk_prepass", "judge")


def _run(tag, rp, ci, w, iters, profile, **kw):
    n = len(rp) - 1
    init = graphs.uniform_init(n, 0)
    with api.Graph(rp, ci, w) as g:
        g.run(init, iters=1, **kw)  # warm-up
        _, st = g.run(init, iters=iters, profile=int(profile), **kw)
    rec = dict(workload=tag, n=n, nnz=int(len(ci)), iters=iters, seconds=st.seconds,
               vertex_updates_per_s=n * iters / st.seconds, energy=st.energy, params=kw)
    if profile:
        rec["kernel_ms_per_iter"] = dict(zip(CLASSES, [m / iters for m in st.kernel_ms]))
        rec.update(walk=st.walk, walk_lag=st.walk_lag, walk_batches_helped=st.walk_helped, far_cells=st.far_cells)
    return rec


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("name")
    ap.add_argument("--iters", type=int, default=5)
    ap.add_argument("--profile", action="store_true")
    ap.add_argument("--precision", type=int, default=api.PREC_MIXED)
    ap.add_argument("--calib-copy", action="store_true",
                    help="first copy 256 MiB device to device (the reference for scaling HBM counters)")
    a = ap.parse_args(argv)
    if a.calib_copy:
        import torch

        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")  # 256 MiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    t0 = time.time()
    out = []
    if a.name == "grid64-algo2":
        rp, ci = graphs.csr_from_edges(*graphs.grid(64, 64))
        out.append(_run(a.name, rp, ci, None, a.iters, a.profile, algo=2, batch=256, precision=a.precision))
    elif a.name in ("rgg250k-algo2", "rgg250k-algo4"):
        rp, ci = graphs.csr_from_edges(*graphs.rgg(250_000, seed=1))
        algo = 2 if a.name.endswith("2") else 4
        out.append(_run(a.name, rp, ci, None, a.iters, a.profile, algo=algo, batch=256, bht=1.2,
                        precision=a.precision))
    elif a.name in ("rmat20-algo4", "rmat20-algo8"):
        n, e, w = graphs.rmat(20, 8, seed=2)
        rp, ci, wc = graphs.csr_from_edges(n, e, w)
        algo = 4 if a.name.endswith("4") else 8
        out.append(_run(a.name, rp, ci, wc, a.iters, a.profile, algo=algo, batch=1024, bht=1.2,
                        precision=a.precision))
    elif a.name == "delaunay2m-algo40":
        rp, ci = graphs.csr_from_edges(*graphs.delaunay_holes())
        out.append(_run(a.name, rp, ci, None, a.iters, a.profile, algo=api.ALGO_FAR_MONO, batch=256, bht=0.4,
                        precision=a.precision))
    elif a.name == "delaunay2m-far":
        from . import farmodel

        model, info = farmodel.train(theta=1.0)
        print(json.dumps(dict(workload=a.name, model_training=info)), flush=True)
        rp, ci = graphs.csr_from_edges(*graphs.delaunay_holes())
        n = len(rp) - 1
        init = graphs.uniform_init(n, 0)
        with api.Graph(rp, ci) as g:
            g.set_far_model(model)
            for algo, mode in ((api.ALGO_FAR_MONO, "monopole"), (api.ALGO_FAR_QUAD, "quadrupole"),
                               (api.ALGO_FAR_LEARNED, "learned"), (api.ALGO_BH, "reference_bh_unit_vector")):
                g.run(init, algo=algo, iters=1, batch=256, bht=1.0)
                _, st = g.run(init, algo=algo, iters=a.iters, batch=256, bht=1.0, profile=int(a.profile))
                rec = dict(workload=a.name, mode=mode, n=n, nnz=int(len(ci)), iters=a.iters, seconds=st.seconds,
                           vertex_updates_per_s=n * a.iters / st.seconds, energy=st.energy)
                if a.profile:
                    rec["kernel_ms_per_iter"] = dict(zip(CLASSES, [m / a.iters for m in st.kernel_ms]))
                    rec["far_cells_per_iter"] = st.far_cells
                out.append(rec)
                print(json.dumps(rec), flush=True)
        out = []
    elif a.name == "batched4096":
        gl = graphs.batched_small(4096, seed=4)
        gs = [graphs.csr_from_edges(nv, e) for nv, e in gl]
        xys = [graphs.uniform_init(len(rp) - 1, k) for k, (rp, _) in enumerate(gs)]
        N = sum(len(rp) - 1 for rp, _ in gs)
        _, _, fin, st = api.run_batched(gs, xys, algo=2, iters=a.iters, batch=128, precision=a.precision)
        out.append(dict(workload=a.name, graphs=len(gs), vertices=N, iters=a.iters, seconds=st.seconds,
                        graphs_per_s=len(gs) / st.seconds, vertex_updates_per_s=N * a.iters / st.seconds,
                        finish_ms_p50=float(np.percentile(fin, 50)), finish_ms_p99=float(np.percentile(fin, 99))))
    elif a.name == "batched4096-pergraph":
        import torch

        gl = graphs.batched_small(4096, seed=4)
        total, N, t_prog = 0.0, 0, time.time()
        for k, (nv, e) in enumerate(gl):
            rp, ci = graphs.csr_from_edges(nv, e)
            xy = torch.from_numpy(graphs.uniform_init(len(rp) - 1, k)).cuda()
            torch.cuda.synchronize()
            with api.Graph(rp, ci) as g:
                st = g.run_device(xy.data_ptr(), algo=2, iters=a.iters, batch=128, precision=a.precision)
            total += st.seconds
            N += len(rp) - 1
            if time.time() - t_prog > 20:
                print(f"# {k + 1} graphs, {total:.1f} s of layout so far", file=sys.stderr, flush=True)
                t_prog = time.time()
        out.append(dict(workload=a.name, graphs=len(gl), vertices=N, iters=a.iters, seconds=total,
                        graphs_per_s=len(gl) / total, vertex_updates_per_s=N * a.iters / total,
                        note="sum of bl_run_device wall times (launch to synchronize), graph set-up excluded"))
    else:
        ap.error(f"unknown workload {a.name}")
    for rec in out:
        print(json.dumps(rec), flush=True)
    print(f"# {a.name}: {time.time() - t0:.1f} s", file=sys.stderr)


if __name__ == "__main__":
    main()
