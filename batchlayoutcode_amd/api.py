"""ctypes binding of libbatchlayout.so (the C ABI in include/batchlayout.h).

The library is the product: every layout, Morton, tree and force call below runs the
HIP kernels for gfx950.  There is no CPU fallback; a missing library or device raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libbatchlayout.so")

BL_OK = 0
STATUS = {0: "BL_OK", -1: "BL_E_ARG", -2: "BL_E_IO", -3: "BL_E_FORMAT", -4: "BL_E_DEVICE",
          -5: "BL_E_DUPKEY", -6: "BL_E_NOMEM", -7: "BL_E_GRAPH", -8: "BL_E_SCRATCH"}
ALGO_CACHE, ALGO_LINLOG, ALGO_BH, ALGO_BHCACHE, ALGO_FA = 2, 3, 4, 6, 8
# physically consistent Barnes-Hut layouts (no reference counterpart)
ALGO_FAR_MONO, ALGO_FAR_EXACT, ALGO_FAR_LEARNED, ALGO_FAR_QUAD = 40, 41, 42, 43
FAR_MONO, FAR_EXACT, FAR_LEARNED, FAR_QUAD = 1, 2, 3, 4
FAR_MODEL_PARAMS, FAR_FEATURES = 9026, 8
PREC_MIXED, PREC_FP64, PREC_FP32 = 0, 1, 2
BATCHED_MAX_N = 8192
# the Barnes-Hut family's batch walks (Stats.walk)
WALK_NONE, WALK_ROW_ORDER, WALK_PF, WALK_SPLIT, WALK_ROW, WALK_GRID, WALK_LAG, WALK_WIDE = 0, 1, 2, 3, 4, 5, 6, 7


class BLError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what} failed: {STATUS.get(code, code)}")
        self.code = code


class bl_params(C.Structure):
    _fields_ = [("algo", C.c_int32), ("iters", C.c_int32), ("batch", C.c_int32), ("converged", C.c_int32),
                ("bht", C.c_double), ("engt", C.c_double), ("weight", C.c_double), ("step0", C.c_double),
                ("precision", C.c_int32), ("profile", C.c_int32)]


class bl_stats(C.Structure):
    _fields_ = [("energy", C.c_double), ("iters_done", C.c_int32), ("status", C.c_int32),
                ("seconds", C.c_double), ("step_final", C.c_double), ("kernel_ms", C.c_double * 8),
                ("kernel_launches", C.c_int32 * 8), ("walk", C.c_int32), ("far_cells", C.c_int64),
                ("far_model_iters", C.c_int32), ("far_judge_err", C.c_float * 2), ("walk_helped", C.c_int32),
                ("walk_lag", C.c_int32), ("nonfinite", C.c_int32)]


_lib = None
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")


def lib():
    """Loads libbatchlayout.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "or `make -C batchlayoutcode_amd/csrc`")
        L = C.CDLL(LIB_PATH)
        L.bl_params_default.argtypes = [C.POINTER(bl_params)]
        L.bl_graph_create.argtypes = [C.c_int32, C.c_int64, _i32p, C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]
        L.bl_graph_destroy.argtypes = [C.c_void_p]
        L.bl_graph_size.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int64)]
        L.bl_run.argtypes = [C.c_void_p, _f64p, C.POINTER(bl_params), C.POINTER(bl_stats)]
        L.bl_run_device.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(bl_params), C.POINTER(bl_stats), C.c_void_p]
        L.bl_run_device_static.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(bl_params), C.POINTER(bl_stats),
                                           C.c_void_p]
        L.bl_mtx_read.argtypes = [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                                  C.POINTER(C.POINTER(C.c_int32)), C.POINTER(C.POINTER(C.c_int32))]
        L.bl_free.argtypes = [C.c_void_p]
        L.bl_initf_read.argtypes = [C.c_char_p, C.c_int32, _f64p, C.POINTER(C.c_int32)]
        L.bl_output_write.argtypes = [C.c_char_p, C.c_int32, _f64p]
        L.bl_bh_morton.argtypes = [C.c_int32, _f64p, _u32p, _u32p, C.POINTER(C.c_double)]
        L.bl_bh_tree.argtypes = [C.c_int32, _f64p, _i32p, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                 _f64p, _f64p, _f64p, _i32p]
        L.bl_bh_forces.argtypes = [C.c_int32, _f64p, C.c_double, C.c_int32, _f64p, _f64p, C.c_int32]
        L.bl_far_forces.argtypes = [C.c_int32, _f64p, C.c_double, C.c_int32, C.c_void_p, C.c_int32, _f64p, _f64p]
        L.bl_far_samples.argtypes = [C.c_int32, _f64p, C.c_double, C.c_int32, _f64p, C.c_int32, _f32p, _f32p, _i32p]
        L.bl_far_model_init.argtypes = [_f32p, C.c_uint64]
        L.bl_far_model_eval.argtypes = [_f32p, C.c_int32, _f32p, _f32p]
        L.bl_far_model_grad.argtypes = [_f32p, C.c_int32, _f32p, _f32p, C.POINTER(C.c_float), _f32p]
        L.bl_far_model_train.argtypes = [_f32p, C.c_int32, _f32p, _f32p, C.c_int32, C.c_int32, C.c_float, C.c_uint64,
                                         _f32p]
        L.bl_graph_set_far_model.argtypes = [C.c_void_p, C.c_void_p]
        L.bl_run_batched.argtypes = [C.c_int32, _i32p, _i64p, _i32p, _i32p, _f64p, C.POINTER(bl_params), _f64p,
                                     _f64p, C.POINTER(bl_stats)]
        L.bl_device_info.argtypes = [C.c_char_p, C.c_int32]
        _lib = L
    return _lib


def _chk(rc: int, what: str):
    if rc != BL_OK:
        raise BLError(rc, what)


def params(**kw) -> bl_params:
    p = bl_params()
    lib().bl_params_default(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise TypeError(f"unknown parameter {k}")
        setattr(p, k, v)
    return p


@dataclass
class Stats:
    energy: float
    iters_done: int
    seconds: float
    step_final: float
    kernel_ms: tuple = (0.0,) * 8
    kernel_launches: tuple = (0,) * 8
    walk: int = 0
    far_cells: int = 0
    far_model_iters: int = 0
    far_judge_err: tuple = (0.0, 0.0)
    walk_helped: int = 0
    walk_lag: int = 0
    nonfinite: int = 0


def _stats(st: bl_stats) -> Stats:
    return Stats(st.energy, st.iters_done, st.seconds, st.step_final, tuple(st.kernel_ms),
                 tuple(st.kernel_launches), st.walk, st.far_cells, st.far_model_iters, tuple(st.far_judge_err),
                 st.walk_helped, st.walk_lag, st.nonfinite)


class Graph:
    """Device-resident CSR graph (bl_graph_create); frees its device memory on close()."""

    def __init__(self, rowptr, colids, w_edge=None):
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        self.colids = np.ascontiguousarray(colids, dtype=np.int32)
        self.w_edge = None if w_edge is None else np.ascontiguousarray(w_edge, dtype=np.float32)
        self.n = len(self.rowptr) - 1
        self.nnz = len(self.colids)
        h = C.c_void_p()
        _chk(lib().bl_graph_create(self.n, self.nnz, self.rowptr,
                                   self.colids.ctypes.data if self.nnz else None,
                                   None if self.w_edge is None else self.w_edge.ctypes.data, C.byref(h)),
             "bl_graph_create")
        self._h = h

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            lib().bl_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def run(self, xy, **kw):
        """Layout from host positions; returns (new positions, Stats)."""
        out = np.ascontiguousarray(np.array(xy, dtype=np.float64).reshape(self.n, 2))
        p = params(**kw)
        st = bl_stats()
        _chk(lib().bl_run(self._h, out, C.byref(p), C.byref(st)), "bl_run")
        return out, _stats(st)

    def set_far_model(self, params):
        """Attaches the learned far-field model used by ALGO_FAR_LEARNED (None detaches)."""
        m = None if params is None else np.ascontiguousarray(params, np.float32)
        _chk(lib().bl_graph_set_far_model(self._h, None if m is None else m.ctypes.data), "bl_graph_set_far_model")

    def run_device(self, d_xy_ptr: int, stream: int | None = None, **kw) -> Stats:
        """Layout on device-resident positions (a raw device pointer, e.g. tensor.data_ptr())."""
        p = params(**kw)
        st = bl_stats()
        _chk(lib().bl_run_device(self._h, C.c_void_p(d_xy_ptr), C.byref(p), C.byref(st),
                                 None if stream is None else C.c_void_p(stream)), "bl_run_device")
        return _stats(st)

    def run_device_static(self, d_xy_ptr: int, stream: int | None = None, **kw) -> Stats:
        """run_device for algos 2, 3 and 8 with the workspace in the library's static device
        array: the call allocates no device memory (BLError BL_E_SCRATCH when it does not fit)."""
        p = params(**kw)
        st = bl_stats()
        _chk(lib().bl_run_device_static(self._h, C.c_void_p(d_xy_ptr), C.byref(p), C.byref(st),
                                        None if stream is None else C.c_void_p(stream)), "bl_run_device_static")
        return _stats(st)


def mtx_read(path: str):
    L = lib()
    n, nnz = C.c_int32(), C.c_int64()
    rp, ci = C.POINTER(C.c_int32)(), C.POINTER(C.c_int32)()
    _chk(L.bl_mtx_read(path.encode(), C.byref(n), C.byref(nnz), C.byref(rp), C.byref(ci)), "bl_mtx_read")
    rowptr = np.ctypeslib.as_array(rp, shape=(n.value + 1,)).copy()
    colids = np.ctypeslib.as_array(ci, shape=(max(nnz.value, 1),))[: nnz.value].copy()
    L.bl_free(C.cast(rp, C.c_void_p))
    L.bl_free(C.cast(ci, C.c_void_p))
    return rowptr, colids


def initf_read(path: str, n: int):
    xy = np.zeros((n, 2))
    k = C.c_int32()
    _chk(lib().bl_initf_read(path.encode(), n, xy, C.byref(k)), "bl_initf_read")
    return xy, k.value


def output_write(path: str, xy) -> None:
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    _chk(lib().bl_output_write(path.encode(), len(xy), xy), "bl_output_write")


def bh_morton(xy):
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    n = len(xy)
    xs, ys = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
    mm = C.c_double()
    _chk(lib().bl_bh_morton(n, xy, xs, ys, C.byref(mm)), "bl_bh_morton")
    return xs, ys, mm.value


def bh_tree(xy):
    """Compressed quadtree: dict(perm, root, diam, cx, cy, child[m,4])."""
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    n = len(xy)
    cap = max(n - 1, 1)
    perm = np.zeros(n, np.int32)
    nn, root = C.c_int32(), C.c_int32()
    dia, cx, cy = np.zeros(cap), np.zeros(cap), np.zeros(cap)
    ch = np.zeros(4 * cap, np.int32)
    _chk(lib().bl_bh_tree(n, xy, perm, cap, C.byref(nn), C.byref(root), dia, cx, cy, ch), "bl_bh_tree")
    m = nn.value
    return dict(perm=perm, root=root.value, diam=dia[:m], cx=cx[:m], cy=cy[:m], child=ch[: 4 * m].reshape(m, 4))


def bh_forces(xy, theta, queries=None, precision=PREC_FP32):
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    q = xy if queries is None else np.ascontiguousarray(queries, dtype=np.float64)
    out = np.zeros((len(q), 2))
    _chk(lib().bl_bh_forces(len(xy), xy, theta, len(q), q, out, precision), "bl_bh_forces")
    return out


def far_forces(xy, theta, mode=FAR_MONO, queries=None, model=None):
    """Repulsion sum_j (x_j - q)/|x_j - q|^2: Barnes-Hut with monopoles (FAR_MONO), second-order
    multipoles (FAR_QUAD), the learned correction (FAR_LEARNED, needs model) or exactly."""
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    q = xy if queries is None else np.ascontiguousarray(queries, dtype=np.float64)
    out = np.zeros((len(q), 2))
    m = None if model is None else np.ascontiguousarray(model, dtype=np.float32)
    _chk(lib().bl_far_forces(len(xy), xy, theta, mode, None if m is None else m.ctypes.data, len(q), q, out),
         "bl_far_forces")
    return out


def far_samples(xy, theta, queries, cap=512):
    """Training rows of the accepted cells of each query: (features[k, 8], targets[k, 2]) with
    the exact (a, b), concatenated in query order."""
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    q = np.ascontiguousarray(queries, dtype=np.float64)
    feat = np.zeros((len(q) * cap, FAR_FEATURES), np.float32)
    targ = np.zeros((len(q) * cap, 2), np.float32)
    cnt = np.zeros(len(q), np.int32)
    _chk(lib().bl_far_samples(len(xy), xy, theta, len(q), q, cap, feat, targ, cnt), "bl_far_samples")
    keep = (np.arange(cap)[None, :] < cnt[:, None]).ravel()
    return feat[keep], targ[keep]


def far_model_init(seed=1):
    p = np.zeros(FAR_MODEL_PARAMS, np.float32)
    _chk(lib().bl_far_model_init(p, seed), "bl_far_model_init")
    return p


def far_model_eval(params, feat):
    feat = np.ascontiguousarray(feat, np.float32)
    out = np.zeros((len(feat), 2), np.float32)
    _chk(lib().bl_far_model_eval(np.ascontiguousarray(params, np.float32), len(feat), feat, out), "bl_far_model_eval")
    return out


def far_model_grad(params, feat, targ):
    feat = np.ascontiguousarray(feat, np.float32)
    targ = np.ascontiguousarray(targ, np.float32)
    g = np.zeros(FAR_MODEL_PARAMS, np.float32)
    loss = C.c_float(0)
    _chk(lib().bl_far_model_grad(np.ascontiguousarray(params, np.float32), len(feat), feat, targ, C.byref(loss), g),
         "bl_far_model_grad")
    return loss.value, g


def far_model_train(params, feat, targ, epochs, batch=4096, lr=1e-3, seed=1):
    p = np.array(params, np.float32, copy=True)
    feat = np.ascontiguousarray(feat, np.float32)
    targ = np.ascontiguousarray(targ, np.float32)
    losses = np.zeros(max(epochs, 1), np.float32)
    _chk(lib().bl_far_model_train(p, len(feat), feat, targ, epochs, batch, lr, seed, losses), "bl_far_model_train")
    return p, losses[:epochs]


def run_batched(graphs, xys, **kw):
    """Lays out a list of (rowptr, colids) graphs in one launch.

    Returns (list of xy, energies, finish_ms, Stats); finish_ms[k] is graph k's completion
    time measured from the start of the launch."""
    count = len(graphs)
    voff = np.zeros(count + 1, np.int32)
    eoff = np.zeros(count + 1, np.int64)
    for k, (rp, ci) in enumerate(graphs):
        voff[k + 1] = voff[k] + len(rp) - 1
        eoff[k + 1] = eoff[k] + len(ci)
    rowptr = np.zeros(voff[-1] + 1, np.int32)
    colids = np.zeros(max(int(eoff[-1]), 1), np.int32)
    for k, (rp, ci) in enumerate(graphs):
        rowptr[voff[k]:voff[k + 1] + 1] = np.asarray(rp, np.int64) + eoff[k]
        colids[eoff[k]:eoff[k + 1]] = ci
    xy = np.ascontiguousarray(np.concatenate([np.asarray(x, np.float64).reshape(-1, 2) for x in xys]))
    energies = np.zeros(count)
    finish = np.zeros(count)
    p = params(**kw)
    st = bl_stats()
    _chk(lib().bl_run_batched(count, voff, eoff, rowptr, colids, xy, C.byref(p), energies, finish, C.byref(st)),
         "bl_run_batched")
    out = [xy[voff[k]:voff[k + 1]].copy() for k in range(count)]
    return out, energies, finish, _stats(st)


def device_info() -> str:
    buf = C.create_string_buffer(256)
    lib().bl_device_info(buf, 256)
    return buf.value.decode()
