"""Training of the learned far-field model on cells from the method's own layouts.

Point sets come from layouts (physically consistent Barnes-Hut, algo 40) of small synthetic
graphs; every query's accepted cells at the opening threshold theta give (features, exact
correction) rows, produced on the device by bl_far_samples; the model is trained on the
device by bl_far_model_train.  The model predicts (a, b) / (side/rho)^2, so the targets are
scaled by the eighth feature.
"""
import time

import numpy as np

from . import api, graphs


def training_cells(theta=1.0, seed=12, sizes=(140, 20_000, 14), layout_iters=15, queries=3000):
    """Rows from three layouts: a grid (sizes[0]^2), a random geometric graph (sizes[1]) and an
    R-MAT graph (scale sizes[2])."""
    rng = np.random.default_rng(seed)
    sources = [graphs.grid(sizes[0], sizes[0]), graphs.rgg(sizes[1], seed=seed + 9),
               graphs.rmat(sizes[2], 8, seed=seed + 10)[:2]]
    feats, targs = [], []
    for k, (nv, e) in enumerate(sources):
        rp, ci = graphs.csr_from_edges(nv, e)
        with api.Graph(rp, ci) as g:
            xy, _ = g.run(graphs.uniform_init(nv, 30 + k), algo=api.ALGO_FAR_MONO, iters=layout_iters, batch=256,
                          bht=theta)
        q = xy[rng.choice(nv, min(nv, queries), replace=False)]
        f, t = api.far_samples(xy, theta, q, cap=512)
        feats.append(f)
        targs.append(t / f[:, 7:8])
    return np.concatenate(feats), np.concatenate(targs)


def train(theta=1.0, epochs=12, batch=4096, lr=2e-3, seed=7):
    """Returns (parameters, info dict).  Cells are drawn at the opening threshold theta, so the
    model covers every threshold up to it."""
    t0 = time.time()
    F, T = training_cells(theta)
    t1 = time.time()
    model, losses = api.far_model_train(api.far_model_init(1), F, T, epochs=epochs, batch=batch, lr=lr, seed=seed)
    info = dict(theta=theta, training_cells=int(len(F)), epochs=epochs, batch=batch, lr=lr,
                loss_first=float(losses[0]), loss_last=float(losses[-1]), data_seconds=t1 - t0,
                train_seconds=time.time() - t1)
    return model, info
