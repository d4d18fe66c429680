"""ctypes binding of the CPU oracle (oracle/build/liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg; never by the
product package.  All arrays are numpy; positions are ``float64[n, 2]``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")
_lib = None

LAW_FR, LAW_LINLOG, LAW_FA = 2, 3, 8
EDUPKEY = -3

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_mtx_read.argtypes = [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                                   C.POINTER(C.POINTER(C.c_int32)), C.POINTER(C.POINTER(C.c_int32))]
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_minibatch.argtypes = [C.c_int, C.c_int32, _i32p, _i32p, C.c_void_p, _f64p, C.c_int, C.c_int,
                                    C.c_double, C.c_double, C.c_double, C.c_int, C.c_double, C.c_int,
                                    C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.orc_bh_layout.argtypes = [C.c_int32, _i32p, _i32p, C.c_void_p, _f64p, C.c_int, C.c_int, C.c_double,
                                    C.c_double, C.c_double, C.c_int, C.POINTER(C.c_double)]
        L.orc_hybrid_layout.argtypes = [C.c_int32, _i32p, _i32p, _f64p, C.c_int, C.c_int, C.c_double, C.c_int,
                                        C.POINTER(C.c_double)]
        L.orc_morton.argtypes = [C.c_int32, _f64p, _u32p, _u32p, C.POINTER(C.c_double)]
        L.orc_bh_tree.argtypes = [C.c_int32, _f64p, _i32p, C.c_int32, C.POINTER(C.c_int32), _i32p, _i32p,
                                  _f64p, _f64p, _f64p, _i32p]
        L.orc_bh_forces.argtypes = [C.c_int32, _f64p, C.c_double, C.c_int32, _f64p, _f64p, C.c_int]
        L.orc_exact_repulsion.argtypes = [C.c_int32, _f64p, C.c_int32, _i32p, _f64p, C.c_int]
        L.orc_far_forces.argtypes = [C.c_int32, _f64p, C.c_double, C.c_int32, _f64p, _f64p, C.c_int]
        L.orc_far_samples.argtypes = [C.c_int32, _f64p, C.c_double, C.c_double, C.c_double, C.c_int32, _f64p,
                                      _f64p, _i32p, C.POINTER(C.c_int32), _f64p]
        L.orc_far_layout.argtypes = [C.c_int32, _i32p, _i32p, C.c_void_p, _f64p, C.c_int, C.c_int, C.c_double,
                                     C.c_double, C.c_double, C.c_int, C.POINTER(C.c_double)]
        _lib = L
    return _lib


def _chk(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with status {rc}")


def mtx_read(path: str):
    L = lib()
    n, nnz = C.c_int32(), C.c_int64()
    rp, ci = C.POINTER(C.c_int32)(), C.POINTER(C.c_int32)()
    _chk(L.orc_mtx_read(path.encode(), C.byref(n), C.byref(nnz), C.byref(rp), C.byref(ci)), "mtx_read")
    rowptr = np.ctypeslib.as_array(rp, shape=(n.value + 1,)).copy()
    colids = np.ctypeslib.as_array(ci, shape=(max(nnz.value, 1),))[: nnz.value].copy()
    L.orc_free(C.cast(rp, C.c_void_p))
    L.orc_free(C.cast(ci, C.c_void_p))
    return rowptr, colids


def _wptr(w):
    return None if w is None else np.ascontiguousarray(w, dtype=np.float32).ctypes.data


def minibatch(rowptr, colids, xy, iters, batch, law=LAW_FR, weight=1.0, step0=1.0, decay=None,
              converged=False, engt=0.01, threads=0, w_edge=None):
    """Runs the minibatch family in place on a copy of xy; returns (xy, energy, iterations)."""
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    colids = np.ascontiguousarray(colids, np.int32)
    out = np.ascontiguousarray(np.array(xy, dtype=np.float64))
    if decay is None:
        decay = 0.999 if law == LAW_FR else 0.99
    w = None if w_edge is None else np.ascontiguousarray(w_edge, np.float32)
    e, k = C.c_double(), C.c_int()
    _chk(lib().orc_minibatch(law, len(rowptr) - 1, rowptr, colids, None if w is None else w.ctypes.data, out,
                             iters, batch, weight, step0, decay, int(converged), engt, threads, C.byref(e),
                             C.byref(k)), "minibatch")
    return out, e.value, k.value


def bh_layout(rowptr, colids, xy, iters, batch, theta=1.2, weight=1.0, step0=1.0, threads=0, w_edge=None):
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    colids = np.ascontiguousarray(colids, np.int32)
    out = np.ascontiguousarray(np.array(xy, dtype=np.float64))
    w = None if w_edge is None else np.ascontiguousarray(w_edge, np.float32)
    e = C.c_double()
    _chk(lib().orc_bh_layout(len(rowptr) - 1, rowptr, colids, None if w is None else w.ctypes.data, out, iters,
                             batch, theta, weight, step0, threads, C.byref(e)), "bh_layout")
    return out, e.value


def hybrid_layout(rowptr, colids, xy, iters, batch, weight=1.0, threads=0):
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    colids = np.ascontiguousarray(colids, np.int32)
    out = np.ascontiguousarray(np.array(xy, dtype=np.float64))
    e = C.c_double()
    _chk(lib().orc_hybrid_layout(len(rowptr) - 1, rowptr, colids, out, iters, batch, weight, threads,
                                 C.byref(e)), "hybrid_layout")
    return out, e.value


def morton(xy):
    xy = np.ascontiguousarray(xy, np.float64)
    n = len(xy)
    xs, ys = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
    mm = C.c_double()
    _chk(lib().orc_morton(n, xy, xs, ys, C.byref(mm)), "morton")
    return xs, ys, mm.value


def bh_tree(xy):
    """Uncompressed quadtree in the reference's pre-order (dict of arrays)."""
    xy = np.ascontiguousarray(xy, np.float64)
    n = len(xy)
    cap = 4 * n + 64
    while True:
        perm = np.zeros(n, np.int32)
        nn = C.c_int32()
        par, qd = np.zeros(cap, np.int32), np.zeros(cap, np.int32)
        dia, cx, cy = np.zeros(cap), np.zeros(cap), np.zeros(cap)
        ch = np.zeros(4 * cap, np.int32)
        rc = lib().orc_bh_tree(n, xy, perm, cap, C.byref(nn), par, qd, dia, cx, cy, ch)
        if rc == -4:
            cap = nn.value + 16
            continue
        _chk(rc, "bh_tree")
        k = nn.value
        return dict(perm=perm, parent=par[:k], quad=qd[:k], diam=dia[:k], cx=cx[:k], cy=cy[:k],
                    child=ch[: 4 * k].reshape(k, 4))


def bh_forces(xy, theta, queries=None, threads=0):
    xy = np.ascontiguousarray(xy, np.float64)
    q = xy if queries is None else np.ascontiguousarray(queries, np.float64)
    out = np.zeros((len(q), 2))
    _chk(lib().orc_bh_forces(len(xy), xy, theta, len(q), q, out, threads), "bh_forces")
    return out


def exact_repulsion(xy, targets, threads=0):
    xy = np.ascontiguousarray(xy, np.float64)
    t = np.ascontiguousarray(targets, np.int32)
    out = np.zeros((len(t), 2))
    _chk(lib().orc_exact_repulsion(len(xy), xy, len(t), t, out, threads), "exact_repulsion")
    return out


def far_forces(xy, theta, queries=None, threads=0):
    """Physically consistent BH (mass-weighted monopole, d/|d|^2 law); theta = 0 is exact."""
    xy = np.ascontiguousarray(xy, np.float64)
    q = xy if queries is None else np.ascontiguousarray(queries, np.float64)
    out = np.zeros((len(q), 2))
    _chk(lib().orc_far_forces(len(xy), xy, theta, len(q), q, out, threads), "far_forces")
    return out


def far_samples(xy, theta, query, cap=100000):
    """Accepted cells of one query: (features[k, 8], targets[k, 2], cells[k, 2], monopole force)."""
    xy = np.ascontiguousarray(xy, np.float64)
    feat = np.zeros((cap, 8))
    targ = np.zeros((cap, 2))
    cell = np.zeros((cap, 2), np.int32)
    cnt = C.c_int32(0)
    force = np.zeros(2)
    _chk(lib().orc_far_samples(len(xy), xy, theta, float(query[0]), float(query[1]), cap, feat, targ, cell,
                               C.byref(cnt), force), "far_samples")
    k = cnt.value
    return feat[:k], targ[:k], cell[:k], force


def far_layout(rowptr, colids, xy, iters, batch, theta, weight=1.0, step0=1.0, w_edge=None, threads=0):
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    colids = np.ascontiguousarray(colids, np.int32)
    xy = np.array(xy, np.float64, copy=True, order="C")
    w = None if w_edge is None else np.ascontiguousarray(w_edge, np.float32)
    e = C.c_double(0)
    _chk(lib().orc_far_layout(len(rowptr) - 1, rowptr, colids, None if w is None else w.ctypes.data, xy, iters,
                              batch, theta, weight, step0, threads, C.byref(e)), "far_layout")
    return xy, e.value
