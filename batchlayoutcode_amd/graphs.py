"""Synthetic graph inputs for the BatchLayout workloads, plus `.mtx` / init-file writers.

Every generator returns ``(n, edges)`` where ``edges`` is an ``int64[m, 2]`` array of
undirected edges ``u < v`` (0-based, no self-loops, no duplicates).  ``csr_from_edges``
turns that into the CSR the reference builds from a ``pattern symmetric`` `.mtx`
(IO.h:60-155 mirrors every off-diagonal entry; CSC.h:147-188 / CSR.h:155-186 leave the
column ids of each row ascending), so the two paths agree entry for entry.

Workloads (sizes as named in the layout benchmarks):
  * ``grid(64, 64)``                       4,096 vertices, 8,064 edges (the CPU-runnable case)
  * ``rgg(250_000, seed=1)``               random geometric graph, mean degree ~8, LCC
  * ``rmat(20, 8, seed=2)``                R-MAT scale 20, (0.57, 0.19, 0.19, 0.05), LCC
  * ``delaunay_holes(2_000_000, seed=3)``  Delaunay edges of points in the unit square minus 20 disks
  * ``batched_small(4096, seed=4)``        many small connected graphs of uneven size
"""
from __future__ import annotations

import numpy as np

__all__ = [
    "grid", "rgg", "rmat", "delaunay_holes", "batched_small", "largest_component",
    "csr_from_edges", "write_mtx", "write_initf", "uniform_init",
]


def _canon(edges: np.ndarray) -> np.ndarray:
    """Sort each pair to (min, max), drop self-loops and duplicates."""
    e = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    e = np.sort(e, axis=1)
    e = e[e[:, 0] != e[:, 1]]
    if len(e) == 0:
        return e
    e = np.unique(e, axis=0)
    return e


def grid(rows: int, cols: int):
    """2-D grid graph: vertex r*cols+c, edges (r,c)-(r,c+1) and (r,c)-(r+1,c)."""
    idx = np.arange(rows * cols, dtype=np.int64).reshape(rows, cols)
    h = np.stack([idx[:, :-1].ravel(), idx[:, 1:].ravel()], axis=1)
    v = np.stack([idx[:-1, :].ravel(), idx[1:, :].ravel()], axis=1)
    return rows * cols, _canon(np.concatenate([h, v]))


def largest_component(n: int, edges: np.ndarray):
    """Keep the largest connected component; relabel vertices 0..n'-1 in id order."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import connected_components

    if len(edges) == 0:
        return 1 if n else 0, edges[:0], np.arange(min(n, 1))
    a = sp.coo_matrix((np.ones(len(edges)), (edges[:, 0], edges[:, 1])), shape=(n, n))
    ncomp, lab = connected_components(a, directed=False)
    big = np.argmax(np.bincount(lab))
    keep = np.flatnonzero(lab == big)
    remap = -np.ones(n, dtype=np.int64)
    remap[keep] = np.arange(len(keep))
    e = remap[edges]
    e = e[(e[:, 0] >= 0) & (e[:, 1] >= 0)]
    return len(keep), _canon(e), keep


def _rgg_edges(pts: np.ndarray, r: float) -> np.ndarray:
    from scipy.spatial import cKDTree

    return np.asarray(cKDTree(pts).query_pairs(r, output_type="ndarray"), dtype=np.int64)


def rgg(n: int = 250_000, seed: int = 1, mean_degree: float = 8.0, return_points: bool = False):
    """Random geometric graph in [0,1)^2 with mean degree mean_degree before the LCC cut.

    The radius is the k-th smallest pair distance, k = mean_degree*n/2, taken from all
    pairs within 1.2*sqrt(d/(pi n)) (boundary vertices have fewer neighbours, so the
    radius lands slightly above the bulk estimate); the largest connected component is kept.
    """
    rng = np.random.default_rng(seed)
    pts = rng.random((n, 2))
    r0 = np.sqrt(mean_degree / (np.pi * n))
    cand = _rgg_edges(pts, 1.2 * r0)
    dist = np.sqrt(((pts[cand[:, 0]] - pts[cand[:, 1]]) ** 2).sum(1))
    k = int(round(mean_degree * n / 2))
    order = np.argsort(dist, kind="stable")[:k]
    r = float(dist[order[-1]])
    e = _canon(cand[order])
    n2, e2, keep = largest_component(n, e)
    if return_points:
        return n2, e2, pts[keep], r
    return n2, e2


def rmat(scale: int = 20, edge_factor: int = 8, a: float = 0.57, b: float = 0.19,
         c: float = 0.19, seed: int = 2, weights: bool = True):
    """R-MAT graph: symmetrised, self-loops and duplicates dropped, LCC kept.

    Returns ``(n, edges, w)`` with FP32 edge weights uniform in [0.5, 2.0] (one per
    undirected edge, same order as ``edges``) when ``weights`` is true.
    """
    rng = np.random.default_rng(seed)
    m = edge_factor << scale
    u = np.zeros(m, dtype=np.int64)
    v = np.zeros(m, dtype=np.int64)
    ab, abc = a + b, a + b + c
    for bit in range(scale):
        r = rng.random(m)
        down = r >= ab            # quadrant c or d: row bit set
        right = ((r >= a) & (r < ab)) | (r >= abc)  # quadrant b or d: col bit set
        u |= down.astype(np.int64) << bit
        v |= right.astype(np.int64) << bit
    n0 = 1 << scale
    e = _canon(np.stack([u, v], axis=1))
    n, e, _ = largest_component(n0, e)
    if not weights:
        return n, e
    w = np.random.default_rng(seed + 1000).uniform(0.5, 2.0, size=len(e)).astype(np.float32)
    return n, e, w


def delaunay_holes(npts: int = 2_000_000, nholes: int = 20, seed: int = 3, return_points: bool = False):
    """Delaunay-triangulation edges of uniform points in [0,1)^2 minus random disks."""
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    pts = rng.random((npts, 2))
    centers = rng.random((nholes, 2))
    radii = rng.uniform(0.02, 0.08, size=nholes)
    keep = np.ones(npts, dtype=bool)
    for cc, rr in zip(centers, radii):
        keep &= ((pts - cc) ** 2).sum(1) > rr * rr
    pts = pts[keep]
    tri = Delaunay(pts).simplices.astype(np.int64)
    e = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [0, 2]]])
    e = _canon(e)
    n, e2, kept = largest_component(len(pts), e)
    if return_points:
        return n, e2, pts[kept]
    return n, e2


def _spanning_tree_plus(nv: int, rng, extra_frac: float = 0.05):
    parent = np.array([rng.integers(0, i) for i in range(1, nv)], dtype=np.int64)
    e = np.stack([parent, np.arange(1, nv, dtype=np.int64)], axis=1)
    k = max(1, int(extra_frac * nv))
    x = rng.integers(0, nv, size=(k, 2))
    return _canon(np.concatenate([e, x]))


def _connected_er(nv: int, rng, mean_degree: float = 6.0):
    # random spanning path guarantees connectivity, plus G(n, p) edges
    perm = rng.permutation(nv)
    path = np.stack([perm[:-1], perm[1:]], axis=1)
    m = int(max(0.0, mean_degree / 2.0 * nv - (nv - 1)))
    x = rng.integers(0, nv, size=(m, 2))
    return _canon(np.concatenate([path, x]))


def batched_small(count: int = 4096, seed: int = 4, nmin: int = 64, nmax: int = 8192):
    """Many small connected graphs, vertex counts log-uniform in [nmin, nmax].

    Mix (round-robin): 2-D grids (closest r x c with r*c <= n), random spanning trees
    plus 5 % extra random edges, and connected Erdos-Renyi with mean degree 6.
    Returns a list of ``(n, edges)``.
    """
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(nmin), np.log(nmax), size=count)).astype(np.int64)
    out = []
    for k, nv in enumerate(sizes):
        kind = k % 3
        if kind == 0:
            r = max(2, int(np.sqrt(nv)))
            c = max(2, int(nv // r))
            out.append(grid(r, c))
        elif kind == 1:
            out.append((int(nv), _spanning_tree_plus(int(nv), rng)))
        else:
            out.append((int(nv), _connected_er(int(nv), rng)))
    return out


def csr_from_edges(n: int, edges: np.ndarray, weights: np.ndarray | None = None):
    """Symmetric CSR (int32 rowptr[n+1], int32 colids[2m]) with ascending colids per row."""
    e = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    rows = np.concatenate([e[:, 0], e[:, 1]])
    cols = np.concatenate([e[:, 1], e[:, 0]])
    order = np.lexsort((cols, rows))
    rows, cols = rows[order], cols[order]
    rowptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(rows, minlength=n), out=rowptr[1:])
    res = [rowptr.astype(np.int32), cols.astype(np.int32)]
    if weights is not None:
        w = np.concatenate([weights, weights]).astype(np.float32)[order]
        res.append(w)
    return tuple(res)


def write_mtx(path: str, n: int, edges: np.ndarray) -> None:
    """`pattern symmetric` MatrixMarket file, one line per undirected edge (row > col)."""
    e = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate pattern symmetric\n")
        f.write(f"{n} {n} {len(e)}\n")
        lines = np.stack([e[:, 1] + 1, e[:, 0] + 1], axis=1)
        np.savetxt(f, lines, fmt="%d %d")


def uniform_init(n: int, seed: int = 0) -> np.ndarray:
    """Initial coordinates uniform in [0,1)^2 from numpy's default generator."""
    return np.random.default_rng(seed).random((n, 2))


def write_initf(path: str, xy: np.ndarray) -> None:
    """`-initf` file: "x y id" per line (read by "%lf %lf %d", algorithms.cpp:112-116)."""
    xy = np.asarray(xy, dtype=np.float64)
    with open(path, "w") as f:
        for i, (x, y) in enumerate(xy):
            f.write("%.17g %.17g %d\n" % (x, y, i + 1))
