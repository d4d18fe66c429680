#!/usr/bin/env python3
"""BatchLayout throughput on one MI355X: learned far field against plain Barnes-Hut.

Workload: the graph of the Delaunay-triangulation edges of 2,000,000 points uniform in the unit
square minus 20 random disks (radius 0.02-0.08), seed 3, largest component (1,662,995 vertices,
9,977,890 CSR entries); initial coordinates uniform in [0,1)^2 (seed 0); the Barnes-Hut iteration
of algo 4 (sample/algorithms.cpp:1020-1083: the tree of the iteration-start positions, attraction
W|d| d over every CSR entry with live positions, sequential batches of -batch 256, STEP *= 0.999)
with the physically consistent repulsion sum_j (x_j - x_i)/|x_j - x_i|^2 (algo 2's law,
:764-768):
  * exact        algo 41, the O(n^2) sum (the FP64 reference the energies are judged against)
  * plain BH     algo 40, mass-weighted monopoles of the cells accepted at -bht theta
  * learned      algo 42, monopoles corrected by the learned far-field model (8-64-64-64-2 GELU,
                 BF16 MFMA, FP32 accumulation), trained first on the device on cells from layouts
                 of small synthetic graphs and random point sets
  * quadrupole   algo 43, monopoles corrected by the analytic second-order term (for comparison)

Qualification: for each mode, theta is swept from large (fast) to small; a theta qualifies when,
after 10 iterations from each of `--starts` uniform starts (seeds 0, 1, ...), the mean over the
starts of the absolute relative difference of the final energy to the exact layout's from the
same start is within 1% (differences of opposite sign do not cancel).  The fastest (largest)
qualifying theta is kept.

A "step" is one layout iteration (every vertex updated once).  `value` = learned vertex-updates/s
= n * steps / seconds of the timed region at the learned mode's qualifying theta, positions
resident in HBM before the region starts; `plain_bh` is the monopole timed the same way at its own
qualifying theta, and `learned_over_plain` their ratio.

Also reported:
  roofline      the dominant kernel of the learned iteration, from per-launch HIP events on the
                library's launch stream (bl_params.profile): the learned traversal against the
                BF16 dense MFMA peak (17,664 flops per evaluated cell), or the batch walk against
                HBM bandwidth (algorithmic bytes per iteration)
  cpu_baseline  the reference built from its own sources (oracle/_ref/ref_fast), algo 4 -bht 1.2,
                timed on this host's cores for a bounded number of iterations of the same graph
  rgg250k_algo2 the all-pairs minibatch step (algo 2, -batch 256, 100 iterations) on the random
                geometric graph of 250,000 points (seed 1, mean degree 8, largest component), with
                the pair kernel's roofline and the reference's algo 2 on the host
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "vertex-updates/s at energy within 1% of FP64 ref; learned vs plain BH, 1 GPU"
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
PEAK_BF16_TFLOPS = 2500.0
PEAK_HBM_GBS = 8000.0
FLOPS_PER_PAIR = 10
FLOPS_PER_CELL = 2 * (8 * 64 + 64 * 64 + 64 * 64 + 64 * 2)  # the model's MFMA work per evaluated cell
QUAL_ITERS = 10


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_reference(n, rp, ci, init, batch, threads, algo, iters, bht=1.2):
    """`iters` iterations of `algo` on the workload's graph on the host cores (bounded sample): the
    reference built from its own sources with its Makefile flags (oracle/_ref/ref_fast); None when
    that binary is absent."""
    ref_bin = os.path.join(ROOT, "oracle", "_ref", "ref_fast")
    if not os.path.exists(ref_bin):
        return None
    from batchlayoutcode_amd import graphs

    with tempfile.TemporaryDirectory() as td:
        mtx = os.path.join(td, "g.mtx")
        initf = os.path.join(td, "init.txt")
        e = np.stack([np.repeat(np.arange(n), np.diff(rp)), ci], 1)
        e = e[e[:, 0] < e[:, 1]]
        graphs.write_mtx(mtx, n, e)
        graphs.write_initf(initf, init)
        t0 = time.time()
        out = subprocess.run([ref_bin, "run", mtx, initf, str(algo), str(iters), str(batch), str(bht), "1.0", "0.01",
                              str(threads)], capture_output=True, text=True, cwd=td, check=True).stdout
        wall = time.time() - t0
    secs = None
    for ln in out.splitlines():
        if "Wall time required:" in ln:
            secs = float(ln.split(":")[-1])
    if secs is None:
        secs = wall
    what = "algo 2" if algo == 2 else f"algo {algo} (-bht {bht})"
    return dict(value=n * iters / secs, unit="vertex-updates/s", cores=threads, kind="reference",
                sample=f"{iters} iteration(s) of {what} on the same {n}-vertex graph, -batch {batch}, "
                       f"-threads {threads} (the reference built from its own sources with its Makefile "
                       f"flags); {secs:.2f} s")


def measured_traffic(name="traffic_k_pairs32_rgg250k.json"):
    """HBM bytes per launch of a kernel from the committed FETCH_SIZE / WRITE_SIZE runs, each
    scaled by what the counter reads for a copy of known size (the file holds the commands)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    t = json.load(open(path))
    fetch = t["fetch_kib_per_dispatch"] * 262144.0 / t["copy_fetch_kib_for_262144_kib"]
    write = t["write_kib_per_dispatch"] * 262144.0 / t["copy_write_kib_for_262144_kib"]
    return (fetch + write) * 1024.0


class Timer:
    """Layout runs on device-resident positions, timed with HIP events around bl_run_device on the
    caller's stream (the library launches on it)."""

    def __init__(self, torch, g, init):
        self.torch = torch
        self.g = g
        self.dev = torch.device("cuda:0")
        self.d_init = torch.from_numpy(np.ascontiguousarray(init)).to(self.dev)
        self.d_xy = torch.empty_like(self.d_init)
        self.stream = torch.cuda.current_stream(self.dev)

    def run(self, **kw):
        torch = self.torch
        self.d_xy.copy_(self.d_init)
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record(self.stream)
        st = self.g.run_device(self.d_xy.data_ptr(), stream=self.stream.cuda_stream, **kw)
        ev1.record(self.stream)
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / 1e3, st


def c4_leg(args, torch, api, graphs):
    from batchlayoutcode_amd import farmodel

    t0 = time.time()
    n, e = graphs.delaunay_holes()
    rp, ci = graphs.csr_from_edges(n, e)
    init = graphs.uniform_init(n, 0)
    log(f"delaunay2m: n={n} nnz={len(ci)} ({time.time() - t0:.1f} s); device: {api.device_info()}")
    model, info = farmodel.train(theta=1.0)
    log(f"far-field model: {info['training_cells']} training cells, loss {info['loss_first']:.3e} -> "
        f"{info['loss_last']:.3e} ({info['data_seconds'] + info['train_seconds']:.1f} s)")
    starts = [init] + [graphs.uniform_init(n, s) for s in range(1, args.starts)]
    K, W, B = args.steps, args.warmup, 256
    sweeps = {"monopole": (api.ALGO_FAR_MONO, (1.0, 0.8, 0.6, 0.5, 0.4, 0.35, 0.3)),
              "learned": (api.ALGO_FAR_LEARNED, (1.2, 1.0, 0.8, 0.6)),
              "quadrupole": (api.ALGO_FAR_QUAD, (1.2, 1.0, 0.8, 0.6))}
    out = dict(workload="delaunay2m-far", n=n, nnz=int(len(ci)), batch=B, qualification_iterations=QUAL_ITERS,
               starts=args.starts, training=info, model="8-64-64-64-2 GELU, BF16 MFMA, FP32 accumulation")
    with api.Graph(rp, ci) as g:
        g.set_far_model(model)
        tm = Timer(torch, g, init)
        g.run(init, algo=api.ALGO_FAR_EXACT, iters=1, batch=B)  # warm-up
        t1 = time.time()
        e_exact = [g.run(x, algo=api.ALGO_FAR_EXACT, iters=QUAL_ITERS, batch=B)[1].energy for x in starts]
        log(f"exact (algo 41): {QUAL_ITERS} iterations x {len(starts)} starts in {time.time() - t1:.1f} s, "
            f"E {' '.join(f'{x:.6e}' for x in e_exact)}")
        out["exact_energies"] = e_exact
        modes = {}
        for name, (algo, thetas) in sweeps.items():
            sweep, best = [], None
            for th in thetas:
                g.run(init, algo=algo, iters=1, batch=B, bht=th)  # warm-up
                es, secs = [], 0.0
                for x in starts:
                    _, st = g.run(x, algo=algo, iters=QUAL_ITERS, batch=B, bht=th)
                    es.append(st.energy)
                    secs += st.seconds
                per = [(a - b) / b for a, b in zip(es, e_exact)]
                r = dict(theta=th, energy_rel_diff_per_start=per, energy_rel_diff_vs_exact=float(np.mean(np.abs(per))),
                         vertex_updates_per_s=n * QUAL_ITERS * len(starts) / secs)
                sweep.append(r)
                log(f"{name} theta {th}: {r['vertex_updates_per_s']:.3e} v-u/s, E vs exact "
                    f"{r['energy_rel_diff_vs_exact']:.3e} (" + " ".join(f"{p:+.2e}" for p in per) + ")")
                if r["energy_rel_diff_vs_exact"] <= 0.01:
                    best = r
                    break  # smaller theta only costs more
            modes[name] = dict(sweep=sweep, within_1pct=best)
        out["modes"] = modes

        def timed(algo, th):
            for _ in range(W):
                tm.run(algo=algo, iters=1, batch=B, bht=th)
            secs, st = tm.run(algo=algo, iters=K, batch=B, bht=th)
            _, sp = tm.run(algo=algo, iters=K, batch=B, bht=th, profile=1)
            cls = ("traversal", "walk", "energy", "tree", "cell_summaries", "walk_prepass", "judge")
            per_iter = {c: sp.kernel_ms[k] / K for k, c in enumerate(cls)}
            return secs, st, sp, per_iter

        res = {}
        for name, algo in (("learned", api.ALGO_FAR_LEARNED), ("plain_bh", api.ALGO_FAR_MONO)):
            q = modes["learned" if name == "learned" else "monopole"]["within_1pct"]
            if q is None:
                res[name] = None
                log(f"{name}: no theta within 1% of the exact energy")
                continue
            secs, st, sp, per_iter = timed(algo, q["theta"])
            res[name] = dict(theta=q["theta"], value=n * K / secs, ms_per_step=secs * 1e3 / K, energy=st.energy,
                             kernel_ms_per_iter=per_iter, far_cells_per_iter=sp.far_cells, walk=sp.walk,
                             walk_lag=sp.walk_lag, walk_batches_helped=sp.walk_helped,
                             model_iterations=sp.far_model_iters if algo == api.ALGO_FAR_LEARNED else None,
                             judge_median_errors=(dict(learned=sp.far_judge_err[0], analytic=sp.far_judge_err[1])
                                                  if algo == api.ALGO_FAR_LEARNED else None))
            log(f"{name} timed at theta {q['theta']}: {K} iterations in {secs:.3f} s = {n * K / secs:.3e} v-u/s; "
                f"per iteration {json.dumps({k: round(v, 3) for k, v in per_iter.items()})} ms; "
                f"cells {sp.far_cells}")
            if name == "learned":
                res["_profile"] = (sp, per_iter)
    return n, rp, ci, init, out, res


WALK_KERNELS = {1: "k_attract_move", 2: "k_walk_pf", 3: "k_walk_split", 4: "k_late_attract + k_walk_row",
                5: "k_walk_grid", 6: "k_late_attract + k_walk_lag"}


def c4_roofline(n, nnz, lp):
    sp, per_iter = lp
    K = max(1, sp.kernel_launches[0])
    trav_ms = sp.kernel_ms[0] / K
    walk_ms = (sp.kernel_ms[1] + sp.kernel_ms[5]) / max(1, sp.kernel_launches[1])
    if trav_ms >= walk_ms:
        achieved = sp.far_cells * FLOPS_PER_CELL / (trav_ms / 1e3) / 1e12
        return dict(bound="mfma", kernel="k_far_trav<TRAV_LEARNED>", achieved=achieved, peak=PEAK_BF16_TFLOPS,
                    unit="TFLOP/s", frac=achieved / PEAK_BF16_TFLOPS, traffic=None, mean_launch_us=trav_ms * 1e3,
                    launches=K, cells_per_launch=sp.far_cells, flops_per_cell=FLOPS_PER_CELL,
                    note="the model's BF16 MFMA work (2 x (8x64 + 64x64 + 64x64 + 64x2) flops per evaluated cell) "
                         "per launch / mean launch time from HIP events on the launch stream, against MI355X's "
                         "dense BF16 MFMA peak; the traversal's own FP64 work and the GELUs are not counted")
    # the batch walk: attraction over every CSR entry with live positions, then the move
    bytes_per_iter = n * (4 + 16 + 16 + 16 + 8) + nnz * (4 + 16)
    achieved = bytes_per_iter / (walk_ms / 1e3) / 1e9
    traffic = measured_traffic("traffic_k_walk_lag_delaunay2m.json") if sp.walk == 6 else None
    return dict(bound="hbm", kernel=WALK_KERNELS.get(sp.walk, "walk"), achieved=achieved, peak=PEAK_HBM_GBS, unit="GB/s",
                frac=achieved / PEAK_HBM_GBS, traffic=traffic, mean_launch_us=walk_ms * 1e3,
                launches=sp.kernel_launches[1], bytes_per_launch=bytes_per_iter,
                note="algorithmic bytes of one iteration's attraction and move (4 B rowptr + 16 B own position + "
                     "16 B repulsion + 16 B written position + 8 B energy term per vertex, 4 B column id + 16 B "
                     "neighbour position per CSR entry) / mean time of the walk and its pre-pass; traffic: their "
                     "HBM bytes per iteration from FETCH_SIZE and WRITE_SIZE runs scaled by a copy of known size "
                     "(profiles/traffic_k_walk_lag_delaunay2m.json); the walk is bound by its per-step latency "
                     "(one workgroup walks the batches in sequence, in groups of consecutive batches with no "
                     "edge between them), not by bytes")


def algo2_leg(args, torch, api, graphs):
    """The all-pairs minibatch step (algo 2) on the 250k random geometric graph."""
    t0 = time.time()
    n, edges = graphs.rgg(250_000, seed=1)
    rp, ci = graphs.csr_from_edges(n, edges)
    init = graphs.uniform_init(n, 0)
    log(f"rgg250k: n={n} nnz={len(ci)} ({time.time() - t0:.1f} s)")
    K, W, B = args.algo2_steps, args.warmup, 256
    with api.Graph(rp, ci) as g:
        tm = Timer(torch, g, init)
        for prec in (api.PREC_MIXED, api.PREC_FP32):
            if W > 0:
                tm.run(algo=2, iters=W, batch=B, precision=prec)
        secs, st = tm.run(algo=2, iters=K, batch=B, precision=api.PREC_MIXED)
        log(f"algo 2: {K} iterations in {secs:.3f} s, E={st.energy:.6e}")
        _, sp = tm.run(algo=2, iters=K, batch=B, precision=api.PREC_MIXED, profile=1)
        pair_ms = sp.kernel_ms[0] / max(1, sp.kernel_launches[0])
        move_ms = sp.kernel_ms[1] / max(1, sp.kernel_launches[1])
        log(f"algo 2 profiled: pair kernel {pair_ms * 1e3:.2f} us/launch x {sp.kernel_launches[0]}, "
            f"move {move_ms * 1e3:.2f} us/launch")
        secs32, st32 = tm.run(algo=2, iters=K, batch=B, precision=api.PREC_FP32)
        _, sp32 = tm.run(algo=2, iters=K, batch=B, precision=api.PREC_FP32, profile=1)
        pair32_ms = sp32.kernel_ms[0] / max(1, sp32.kernel_launches[0])
        energy = dict(mixed=st.energy, fp32=st32.energy)
        if not args.no_fp64_check:
            t1 = time.time()
            _, s64 = tm.run(algo=2, iters=K, batch=B, precision=api.PREC_FP64)
            rel = abs(st.energy - s64.energy) / s64.energy
            energy.update(fp64_reference_order=s64.energy, rel_diff=rel, within_1pct=bool(rel <= 0.01),
                          fp32_rel_diff=abs(st32.energy - s64.energy) / s64.energy, fp64_seconds=time.time() - t1)
            log(f"algo 2 FP64 check: E64={s64.energy:.6e}, mixed rel diff {rel:.2e}")
    nb = (n + B - 1) // B
    flops_per_launch = FLOPS_PER_PAIR * n * (n - 1) / nb
    achieved = flops_per_launch / (pair_ms / 1e3) / 1e12
    achieved32 = flops_per_launch / (pair32_ms / 1e3) / 1e12
    out = dict(workload="rgg250k-algo2", n=n, nnz=int(len(ci)), batch=B, iterations=K, value=n * K / secs,
               unit="vertex-updates/s", ms_per_step=secs * 1e3 / K, energy=energy,
               precision="mixed: FP64 differences and sums, FP32-seeded reciprocal with one FP64 Newton step",
               roofline=dict(bound="valu-fp64", kernel="k_pairs32<4,true>", achieved=achieved, peak=PEAK_FP64_TFLOPS,
                             unit="TFLOP/s", frac=achieved / PEAK_FP64_TFLOPS,
                             frac_of_fp32_peak=achieved / PEAK_FP32_TFLOPS, traffic=measured_traffic(),
                             mean_launch_us=pair_ms * 1e3, launches=sp.kernel_launches[0],
                             flops_per_launch=flops_per_launch, move_us=move_ms * 1e3),
               fp32_mode=dict(value=n * K / secs32, ms_per_step=secs32 * 1e3 / K, pair_kernel_us=pair32_ms * 1e3,
                              achieved_tflops=achieved32, frac_of_fp32_peak=achieved32 / PEAK_FP32_TFLOPS))
    if not args.no_cpu:
        threads = os.cpu_count() or 1
        out["cpu_baseline"] = cpu_reference(n, rp, ci, init, B, threads, algo=2, iters=1)
        if out["cpu_baseline"]:
            log(f"algo 2 cpu baseline: {out['cpu_baseline']['value']:.3e} vertex-updates/s")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed layout iterations")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--starts", type=int, default=3, help="uniform starts of the energy qualification")
    ap.add_argument("--algo2-steps", type=int, default=100)
    ap.add_argument("--cpu-iters", type=int, default=5, help="reference algo-4 iterations timed on the host")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-fp64-check", action="store_true")
    ap.add_argument("--no-algo2", action="store_true", help="skip the all-pairs leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU design is described, not built)", file=sys.stderr)
        sys.exit(2)

    import torch
    from batchlayoutcode_amd import api, graphs

    n, rp, ci, init, c4, res = c4_leg(args, torch, api, graphs)
    learned, plain = res.get("learned"), res.get("plain_bh")
    value = learned["value"] if learned else None
    ratio = learned["value"] / plain["value"] if learned and plain else None
    if learned and plain:
        log(f"learned {learned['value']:.3e} v-u/s (theta {learned['theta']}) vs plain BH {plain['value']:.3e} "
            f"(theta {plain['theta']}): {ratio:.3f}x")
    K = args.steps
    out = dict(metric=METRIC, value=value, unit="vertex-updates/s", n_gpus=1, steps=K, warmup=args.warmup,
               ms_per_step=learned["ms_per_step"] if learned else None, higher_is_better=True, scaling="weak",
               vs_baseline=None, dtype="f64",
               data="synthetic (generated graph and uniform starts; the far-field model trained on the device "
                    "on synthetic point sets)",
               config=dict(workload="delaunay2m-far",
                           graph="Delaunay edges of 2,000,000 points in the unit square minus 20 disks (radius "
                                 "0.02-0.08), seed 3, largest component",
                           n=n, nnz=int(len(ci)), batch=256, init="uniform seed 0", parallelism="none (1 GPU)",
                           learned_theta=learned["theta"] if learned else None,
                           plain_bh_theta=plain["theta"] if plain else None,
                           precision="positions, differences and sums FP64; FP32-seeded reciprocals with an FP64 "
                                     "Newton step; the model in BF16 with FP32 accumulation"),
               learned=learned, plain_bh=plain, learned_over_plain=ratio, delaunay2m=c4)
    if "_profile" in res:
        out["roofline"] = c4_roofline(n, len(ci), res["_profile"])
    if not args.no_cpu:
        threads = os.cpu_count() or 1
        log(f"cpu baseline (reference algo 4) on {threads} threads ...")
        out["cpu_baseline"] = cpu_reference(n, rp, ci, init, 256, threads, algo=4, iters=args.cpu_iters)
        if out["cpu_baseline"]:
            log(f"cpu baseline: {out['cpu_baseline']['value']:.3e} vertex-updates/s")
    if not args.no_algo2:
        out["rgg250k_algo2"] = algo2_leg(args, torch, api, graphs)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
