"""hyperalignment_amd -- searchlight Procrustes on MI355X (gfx950).

ctypes binding of ``libhyperalignment_amd.so``. The public function
``searchlight_procrustes`` mirrors the reference wrapper
``hyperalignment_c.py:112-223`` (same arguments, Fortran-order float64 data,
int32 flat searchlight arrays, dense float64 output) with one more backend
name, ``"hip"`` (THaBackend value 4). Everything runs on the GPU; there is no
CPU fallback, and a missing library or device raises.
"""

from __future__ import annotations

import ctypes
import os
import numpy as np

_here = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_here, "libhyperalignment_amd.so")

_lib = None

# THaBackend values (ha_common.h:70-76 plus kHaBackendHIP = 4)
BACKENDS = {"cpu64": 0, "c64": 0, "cpu32": 1, "c32": 1, "metal": 2, "cuda": 3, "hip": 4}
METHODS = {"auto": 0, "jacobi32": 1, "polar": 2, "polar_learned": 3, "jacobi64": 4}
ERRORS = {0: "success", -1: "allocation failed", -2: "SVD failed", -3: "invalid argument",
          -4: "internal / GPU error"}

MLP_IN, MLP_HID, MLP_OUT = 12, 64, 2
MLP_NPARAM = MLP_IN * MLP_HID + MLP_HID + MLP_HID * MLP_HID + MLP_HID + MLP_HID * MLP_OUT + MLP_OUT


class HaError(RuntimeError):
    def __init__(self, fn, rc):
        super().__init__(f"{fn} failed with error code {rc} ({ERRORS.get(rc, 'unknown')})")
        self.rc = rc


class TMat(ctypes.Structure):
    _fields_ = [("data", ctypes.POINTER(ctypes.c_double)), ("rows", ctypes.c_int32),
                ("cols", ctypes.c_int32)]


class TSparseCSC(ctypes.Structure):
    _fields_ = [("data", ctypes.POINTER(ctypes.c_double)), ("indices", ctypes.POINTER(ctypes.c_int32)),
                ("indptr", ctypes.POINTER(ctypes.c_int32)), ("rows", ctypes.c_int32),
                ("cols", ctypes.c_int32), ("nnz", ctypes.c_int64)]


class TSearchlights(ctypes.Structure):
    _fields_ = [("indices", ctypes.POINTER(ctypes.POINTER(ctypes.c_int32))),
                ("dists", ctypes.POINTER(ctypes.POINTER(ctypes.c_double))),
                ("sizes", ctypes.POINTER(ctypes.c_int32)), ("count", ctypes.c_int32),
                ("radius", ctypes.c_double)]


class HaHipOptions(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int32), ("fp64_fallback", ctypes.c_int32),
                ("max_polar_steps", ctypes.c_int32), ("polar_tol", ctypes.c_double),
                ("plain_ell", ctypes.c_double), ("tile", ctypes.c_int32), ("use_graph", ctypes.c_int32),
                ("chunk", ctypes.c_int32), ("verbose", ctypes.c_int32),
                ("learned_margin_scale", ctypes.c_float), ("learned_margin_ell", ctypes.c_float),
                ("team", ctypes.c_int32), ("team_spin", ctypes.c_uint32), ("orth_tol", ctypes.c_double),
                ("template_ell", ctypes.c_double)]


class HaHipStats(ctypes.Structure):
    _fields_ = [("ms_weights", ctypes.c_double), ("ms_crosscov", ctypes.c_double),
                ("ms_polar", ctypes.c_double), ("ms_fallback", ctypes.c_double),
                ("ms_scatter", ctypes.c_double), ("ms_total", ctypes.c_double),
                ("ms_pattern", ctypes.c_double), ("n_fallback", ctypes.c_int64),
                ("n_unconverged", ctypes.c_int64), ("steps_total", ctypes.c_int64),
                ("steps_hist", ctypes.c_int32 * 32), ("nnz", ctypes.c_int64),
                ("gemm_flops", ctypes.c_double), ("kernel_ms_polar_avg", ctypes.c_double),
                ("polar_launches", ctypes.c_int32), ("ms_apply", ctypes.c_double),
                ("ms_corr", ctypes.c_double), ("n_orth_checked", ctypes.c_int64),
                ("n_orth_flagged", ctypes.c_int64), ("orth_max", ctypes.c_double),
                ("n_team_giveup", ctypes.c_int64), ("n_team_handback", ctypes.c_int64),
                ("ms_orth", ctypes.c_double), ("graph_replay", ctypes.c_int32)]

    def as_dict(self):
        d = {name: getattr(self, name) for name, _ in self._fields_}
        d["steps_hist"] = list(self.steps_hist)
        return d


class HaHipTuning(ctypes.Structure):
    _fields_ = [("tile", ctypes.c_int32 * 6), ("team", ctypes.c_int32 * 6), ("sample", ctypes.c_int32 * 6),
                ("searchlights", ctypes.c_int32 * 6), ("ms", (ctypes.c_float * 3) * 6)]


_P = ctypes.POINTER
_d, _i32, _f = ctypes.c_double, ctypes.c_int32, ctypes.c_float
_SIGS = {
    "ha_hip_init": (ctypes.c_int, []),
    "ha_hip_cleanup": (None, []),
    "ha_hip_available": (ctypes.c_bool, []),
    "ha_hip_get_options": (None, [_P(HaHipOptions)]),
    "ha_hip_set_options": (ctypes.c_int, [_P(HaHipOptions)]),
    "ha_hip_set_model": (ctypes.c_int, [_P(_f), _P(_f)]),
    "ha_set_num_threads": (None, [ctypes.c_int]),
    "ha_get_num_threads": (ctypes.c_int, []),
    "ha_searchlight_procrustes_dense": (ctypes.c_int, [_P(_d), _P(_d), _i32, _i32, _P(_i32), _P(_i32), _P(_d),
                                                       _i32, _d, _P(_d), ctypes.c_bool, ctypes.c_bool,
                                                       ctypes.c_int, ctypes.c_bool]),
    "ha_searchlight_procrustes_hip": (ctypes.c_int, [_P(_d), _P(_d), _i32, _i32, _P(_i32), _P(_i32), _P(_d),
                                                     _i32, _d, _P(_d), ctypes.c_bool, ctypes.c_bool,
                                                     ctypes.c_bool]),
    "ha_searchlight_procrustes": (ctypes.c_int, [_P(TMat), _P(TMat), _P(TSearchlights), _P(TSearchlights),
                                                 _P(TSparseCSC), _P(_P(_d)), ctypes.c_bool, ctypes.c_bool,
                                                 ctypes.c_int]),
    "ha_searchlight_weights": (ctypes.c_int, [_P(TSearchlights), _P(_P(_d))]),
    "ha_sparse_init": (_P(TSparseCSC), [_P(TSearchlights), _i32]),
    "ha_sparse_free": (None, [_P(TSparseCSC)]),
    "ha_sparse_scatter_add": (None, [_P(TSparseCSC), _P(_i32), _P(_i32), _i32, _i32, _P(_d), _P(_d)]),
    "ha_procrustes": (ctypes.c_int, [_P(TMat), _P(TMat), _P(TMat), ctypes.c_bool, ctypes.c_bool]),
    "ha_ridge": (ctypes.c_int, [_P(TMat), _P(TMat), _d, _P(TMat)]),
    "ha_polar_newton": (ctypes.c_int, [_P(_f), _P(_f), _i32, ctypes.c_bool, ctypes.c_bool, _P(_f), _i32]),
    "ha_ridge_vec": (ctypes.c_int, [_P(TMat), _P(_d), _d, _P(_d)]),
    "ha_searchlight_ridge": (ctypes.c_int, [_P(TMat), _P(TMat), _P(TSearchlights), _P(TSearchlights),
                                            _P(TSparseCSC), _P(_P(_d)), _d]),
    "ha_hip_plan_ridge": (ctypes.c_int, [ctypes.c_void_p, _d]),
    "ha_mesh_searchlights": (ctypes.c_int, [_P(_d), _i32, _P(_i32), _i32, _d, _P(_i32), _i32, _P(_P(_i32)),
                                            _P(_P(_i32)), _P(_P(_d))]),
    "ha_mesh_searchlights_geodesic": (ctypes.c_int, [_P(_d), _i32, _P(_i32), _i32, _d, _P(_i32), _i32,
                                                     _P(_P(_i32)), _P(_P(_i32)), _P(_P(_d))]),
    "ha_mesh_free": (None, [ctypes.c_void_p]),
    "ha_hip_plan_create": (ctypes.c_int, [_i32, _i32, _P(_i32), _P(_i32), _P(_d), _i32, _d,
                                          _P(ctypes.c_void_p)]),
    "ha_hip_plan_destroy": (None, [ctypes.c_void_p]),
    "ha_hip_plan_load": (ctypes.c_int, [ctypes.c_void_p, _P(_d), _P(_d), ctypes.c_bool]),
    "ha_hip_plan_execute": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_bool, ctypes.c_bool]),
    "ha_hip_plan_nnz": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_int64)]),
    "ha_hip_plan_download": (ctypes.c_int, [ctypes.c_void_p, _P(_i32), _P(_i32), _P(_d)]),
    "ha_hip_plan_download_local": (ctypes.c_int, [ctypes.c_void_p, _P(_i32), _i32, _P(_f)]),
    "ha_hip_plan_stats": (ctypes.c_int, [ctypes.c_void_p, _P(HaHipStats)]),
    "ha_hip_plan_orthogonality": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _P(_d)]),
    "ha_hip_plan_spectra": (ctypes.c_int, [ctypes.c_void_p, _P(_i32), _i32, _P(_f), _P(_f)]),
    "ha_hip_plan_steps": (ctypes.c_int, [ctypes.c_void_p, _P(_i32), _P(_i32)]),
    "ha_hip_plan_predictions": (ctypes.c_int, [ctypes.c_void_p, _P(_f)]),
    "ha_hip_plan_tiles": (ctypes.c_int, [ctypes.c_void_p, _P(_i32)]),
    "ha_hip_plan_tuning": (ctypes.c_int, [ctypes.c_void_p, _P(HaHipTuning)]),
    "ha_hip_plan_det_signs": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _P(_i32)]),
    "ha_hip_plan_evaluate": (ctypes.c_int, [ctypes.c_void_p, _P(_d), _P(_d), _i32, ctypes.c_bool, _P(_d), _P(_d)]),
    "ha_searchlight_template_flat": (ctypes.c_int, [_P(_P(_d)), _i32, _i32, _i32, _P(_i32), _P(_i32), _P(_d), _i32,
                                                     _d, ctypes.c_bool, ctypes.c_bool, ctypes.c_bool, _i32,
                                                     _P(_d), ctypes.c_bool]),
    "ha_template_procrustes": (ctypes.c_int, [_P(_P(TMat)), _i32, ctypes.c_bool, ctypes.c_bool, ctypes.c_bool,
                                              _i32, _P(TMat)]),
    "ha_hip_mlp_forward": (ctypes.c_int, [_P(_f), _P(_f), _P(_f), ctypes.c_int64, _P(_f)]),
    "ha_hip_mlp_grad": (ctypes.c_int, [_P(_f), _P(_f), _P(_f), _P(_f), ctypes.c_int64, _f, _f, _P(_f), _P(_f)]),
    "ha_hip_mlp_train": (ctypes.c_int, [_P(_f), _P(_f), _P(_f), _P(_f), ctypes.c_int64, _i32, _i32, _f, _f, _f,
                                        ctypes.c_uint64, _P(_f)]),
}


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None and a.size != int(np.prod(shape)):
        raise ValueError(f"expected {shape}, got {a.shape}")
    return a


def mlp_forward(params, norm, X):
    """Batch MLP inference on the GPU: (N, 12) features -> (N, 2)."""
    p, n, x = _f32(params, (MLP_NPARAM,)), _f32(norm, (2 * MLP_IN,)), _f32(X)
    y = np.zeros((x.shape[0], MLP_OUT), dtype=np.float32)
    _check("ha_hip_mlp_forward", lib().ha_hip_mlp_forward(_ptr(p, _f), _ptr(n, _f), _ptr(x, _f), x.shape[0],
                                                          _ptr(y, _f)))
    return y


def mlp_grad(params, norm, X, Y, w=(1.0, 1.0)):
    """Full-batch loss gradient on the GPU; returns (grad, loss)."""
    p, n, x, y = _f32(params, (MLP_NPARAM,)), _f32(norm, (2 * MLP_IN,)), _f32(X), _f32(Y)
    g = np.zeros(MLP_NPARAM, dtype=np.float32)
    loss = np.zeros(1, dtype=np.float32)
    _check("ha_hip_mlp_grad", lib().ha_hip_mlp_grad(_ptr(p, _f), _ptr(n, _f), _ptr(x, _f), _ptr(y, _f), x.shape[0],
                                                    float(w[0]), float(w[1]), _ptr(g, _f), _ptr(loss, _f)))
    return g, float(loss[0])


def mlp_train(params, norm, X, Y, epochs=20, batch=4096, lr=1e-3, w=(1.0, 1.0), seed=0):
    """Adam training on the GPU; returns (trained params, per-epoch loss)."""
    p = _f32(params, (MLP_NPARAM,)).copy()
    n, x, y = _f32(norm, (2 * MLP_IN,)), _f32(X), _f32(Y)
    hist = np.zeros(max(epochs, 1), dtype=np.float32)
    _check("ha_hip_mlp_train", lib().ha_hip_mlp_train(_ptr(p, _f), _ptr(n, _f), _ptr(x, _f), _ptr(y, _f),
                                                      x.shape[0], int(epochs), int(batch), float(lr), float(w[0]),
                                                      float(w[1]), int(seed), _ptr(hist, _f)))
    return p, hist[:epochs]


def lib():
    """Loads the built library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} not found: run __graft_entry__.build() (or make in "
                                f"hyperalignment_amd/csrc) first")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct)) if a is not None else None


def _check(fn, rc):
    if rc != 0:
        raise HaError(fn, rc)


def init():
    """ha_hip_init(); raises when no usable GPU is present."""
    _check("ha_hip_init", lib().ha_hip_init())


def available() -> bool:
    return bool(lib().ha_hip_available())


def get_options() -> HaHipOptions:
    o = HaHipOptions()
    lib().ha_hip_get_options(ctypes.byref(o))
    return o


def set_options(**kw):
    """Updates run options, e.g. set_options(method="polar", polar_tol=1e-2)."""
    o = get_options()
    for k, v in kw.items():
        if k == "method" and isinstance(v, str):
            v = METHODS[v]
        setattr(o, k, v)
    _check("ha_hip_set_options", lib().ha_hip_set_options(ctypes.byref(o)))
    return o


_model_loaded = False


def resolved_method(method, kmax):
    """The local solver a pass with this method and largest searchlight runs:
    "auto" picks FP64 Jacobi for k <= 32, else the learned polar iteration when
    a model is loaded, else the plain polar iteration."""
    if method != "auto":
        return method
    if kmax <= 32:
        return "jacobi64"
    return "polar_learned" if _model_loaded else "polar"


def set_model(params, feat_norm):
    """Installs MLP parameters (MLP_NPARAM floats) and feature normalisation (2*MLP_IN)."""
    global _model_loaded
    if params is None:
        _check("ha_hip_set_model", lib().ha_hip_set_model(None, None))
        _model_loaded = False
        return
    p = np.ascontiguousarray(params, dtype=np.float32).ravel()
    n = np.ascontiguousarray(feat_norm, dtype=np.float32).ravel()
    if p.size != MLP_NPARAM or n.size != 2 * MLP_IN:
        raise ValueError("bad MLP parameter sizes")
    _check("ha_hip_set_model", lib().ha_hip_set_model(_ptr(p, ctypes.c_float), _ptr(n, ctypes.c_float)))
    _model_loaded = True


def _flat(sls, dists):
    sizes = np.array([len(s) for s in sls], dtype=np.int32)
    offsets = np.zeros(len(sls) + 1, dtype=np.int32)
    np.cumsum(sizes, out=offsets[1:])
    idx = np.ascontiguousarray(np.concatenate(sls).astype(np.int32))
    d = None if dists is None else np.ascontiguousarray(np.concatenate(dists).astype(np.float64))
    return idx, offsets, d


def searchlight_procrustes(X, Y, sls, dists, radius, backend="hip", reflection=True, scaling=False, n_jobs=1,
                           T=None):
    """Searchlight Procrustes, dense output (hyperalignment_c.py:112-223).

    X, Y: (n_samples, n_vertices). sls / dists: lists of int / float arrays.
    Returns T (n_vertices, n_vertices) float64; if T is given it is
    accumulated in place (the C entry adds into T_out).
    """
    if backend not in BACKENDS:
        raise ValueError(f"backend {backend!r} is not one of {sorted(BACKENDS)}")
    idx, offsets, d = _flat(sls, dists)
    return searchlight_procrustes_flat(X, Y, idx, offsets, d, radius, backend=backend, reflection=reflection,
                                       scaling=scaling, n_jobs=n_jobs, T=T)


def searchlight_procrustes_flat(X, Y, idx, offsets, dists, radius, backend="hip", reflection=True,
                                scaling=False, n_jobs=1, T=None):
    L = lib()
    if n_jobs > 0:
        L.ha_set_num_threads(n_jobs)
    X = np.asfortranarray(X, dtype=np.float64)
    Y = np.asfortranarray(Y, dtype=np.float64)
    nt, nv = X.shape
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    offsets = np.ascontiguousarray(offsets, dtype=np.int32)
    d = None if dists is None else np.ascontiguousarray(dists, dtype=np.float64)
    if T is None:
        T = np.zeros((nv, nv), dtype=np.float64)
    if T.shape != (nv, nv) or T.dtype != np.float64 or not T.flags.c_contiguous:
        raise ValueError("T must be a C-contiguous float64 (nv, nv) array")
    rc = L.ha_searchlight_procrustes_dense(
        _ptr(X, ctypes.c_double), _ptr(Y, ctypes.c_double), nt, nv, _ptr(idx, ctypes.c_int32),
        _ptr(offsets, ctypes.c_int32), _ptr(d, ctypes.c_double), len(offsets) - 1, float(radius),
        _ptr(T, ctypes.c_double), bool(reflection), bool(scaling), BACKENDS[backend], True)
    _check("ha_searchlight_procrustes_dense", rc)
    return T


def searchlight_template_flat(dms, idx, offsets, dists, radius, reflection=True, scaling=False,
                              zscore_common=True, level2_iter=2):
    """Searchlight Procrustes template on the GPU (ha_searchlight_template_flat):
    dms is a sequence of (nt, nv) arrays; returns the (nt, nv) float64 template,
    tpl[:, sl] += local_template(sl) * w as src/hyperalignment/searchlight.py:215-236
    with the Procrustes local template of csrc/ha_template.c:14-114."""
    subs = [np.asfortranarray(d, dtype=np.float64) for d in dms]
    nt, nv = subs[0].shape
    if any(d.shape != (nt, nv) for d in subs):
        raise ValueError("all subjects must have the same shape")
    ptrs = (ctypes.POINTER(ctypes.c_double) * len(subs))(*[_ptr(d, ctypes.c_double) for d in subs])
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    offsets = np.ascontiguousarray(offsets, dtype=np.int32)
    d = None if dists is None else np.ascontiguousarray(dists, dtype=np.float64)
    out = np.zeros((nt, nv), dtype=np.float64, order="F")
    rc = lib().ha_searchlight_template_flat(ptrs, len(subs), nt, nv, _ptr(idx, ctypes.c_int32),
                                            _ptr(offsets, ctypes.c_int32), _ptr(d, ctypes.c_double),
                                            len(offsets) - 1, float(radius), bool(reflection), bool(scaling),
                                            bool(zscore_common), int(level2_iter), _ptr(out, ctypes.c_double),
                                            True)
    _check("ha_searchlight_template_flat", rc)
    return np.ascontiguousarray(out)


def last_template_stats():
    """ha_hip_last_template_stats: the last searchlight_template_flat call's
    kernel time (ms_polar), polar steps (steps_total) and FP64 flops (gemm_flops)."""
    s = HaHipStats()
    _check("ha_hip_last_template_stats", lib().ha_hip_last_template_stats(ctypes.byref(s)))
    return s


def template_procrustes(dms, reflection=True, scaling=False, zscore_common=True, level2_iter=1):
    """ha_template_procrustes on the GPU (csrc/ha_template.h:18-22): the
    Procrustes template of whole (nt, nv) subject matrices."""
    subs = [np.ascontiguousarray(d, dtype=np.float64) for d in dms]
    nt, nv = subs[0].shape
    mats = [TMat(_ptr(d, ctypes.c_double), nt, nv) for d in subs]
    arr = (ctypes.POINTER(TMat) * len(mats))(*[ctypes.pointer(m) for m in mats])
    out = np.zeros((nt, nv), dtype=np.float64)
    tm = TMat(_ptr(out, ctypes.c_double), nt, nv)
    rc = lib().ha_template_procrustes(arr, len(mats), bool(reflection), bool(scaling), bool(zscore_common),
                                      int(level2_iter), ctypes.byref(tm))
    _check("ha_template_procrustes", rc)
    return out


class _SlHolder:
    """Keeps the ctypes arrays of a TSearchlights alive."""

    def __init__(self, sls, dists=None, radius=0.0):
        self.arrs = [np.ascontiguousarray(s, dtype=np.int32) for s in sls]
        self.darrs = None if dists is None else [np.ascontiguousarray(x, dtype=np.float64) for x in dists]
        n = len(self.arrs)
        self.ip = (ctypes.POINTER(ctypes.c_int32) * n)(*[_ptr(a, ctypes.c_int32) for a in self.arrs])
        self.dp = None if self.darrs is None else (ctypes.POINTER(ctypes.c_double) * n)(
            *[_ptr(a, ctypes.c_double) for a in self.darrs])
        self.sizes = np.array([len(a) for a in self.arrs], dtype=np.int32)
        self.st = TSearchlights(self.ip, self.dp, _ptr(self.sizes, ctypes.c_int32), n, float(radius))


def sparse_init(sls, nv=0):
    """ha_sparse_init on the GPU; returns a scipy.sparse.csc_matrix of zeros."""
    import scipy.sparse as sp
    h = _SlHolder(sls)
    m = lib().ha_sparse_init(ctypes.byref(h.st), int(nv))
    if not m:
        raise HaError("ha_sparse_init", -4)
    try:
        c = m.contents
        n, nnz = c.cols, c.nnz
        indptr = np.ctypeslib.as_array(c.indptr, shape=(n + 1,)).copy()
        indices = np.ctypeslib.as_array(c.indices, shape=(max(nnz, 1),))[:nnz].copy()
        return sp.csc_matrix((np.zeros(nnz), indices, indptr), shape=(c.rows, c.cols))
    finally:
        lib().ha_sparse_free(m)


def searchlight_weights(sls, dists=None, radius=0.0):
    """ha_searchlight_weights on the GPU; list of float64 arrays."""
    h = _SlHolder(sls, dists, radius)
    n = len(sls)
    outs = (ctypes.POINTER(ctypes.c_double) * n)()
    _check("ha_searchlight_weights", lib().ha_searchlight_weights(ctypes.byref(h.st), outs))
    libc = ctypes.CDLL(None)
    libc.free.argtypes = [ctypes.c_void_p]
    res = []
    for s in range(n):
        res.append(np.ctypeslib.as_array(outs[s], shape=(len(sls[s]),)).copy())
        libc.free(ctypes.cast(outs[s], ctypes.c_void_p))
    return res


# ---- the Python package's own names (src/hyperalignment/__init__.py:1-7) ----

def compute_searchlight_weights(sls, dists=None, radius=None, return_sparse=False):
    """searchlight.py:12-79 on the GPU: per-searchlight weights summing to one per
    vertex (distance-based when dists is given). return_sparse: a CSR matrix with
    one row per searchlight and one column per vertex."""
    import scipy.sparse as sp
    if dists is not None and radius is None:
        raise ValueError("Radius must be provided if distances are given.")
    w = searchlight_weights(sls, dists, 0.0 if radius is None else radius)
    if not return_sparse:
        return w
    nv = int(max(int(np.max(sl)) for sl in sls)) + 1
    rows = np.concatenate([np.full(len(sl), i, dtype=np.int64) for i, sl in enumerate(sls)])
    cols = np.concatenate([np.asarray(sl, dtype=np.int64) for sl in sls])
    return sp.csr_matrix((np.concatenate(w), (rows, cols)), shape=(len(sls), nv))


def initialize_sparse_matrix(sls, nv=None, dtype=np.float64, cache_fn=None):
    """sparse.py:6-43 with the pattern built on the GPU (ha_sparse_init): a CSC
    matrix of zeros over the union of the searchlights' vertex pairs. cache_fn
    is read with scipy.sparse.load_npz when it exists and written with
    scipy.sparse.save_npz otherwise, so cache files are interchangeable with
    the reference's."""
    import scipy.sparse as sp
    if cache_fn is not None and os.path.exists(cache_fn):
        return sp.load_npz(cache_fn)
    if nv is None:
        nv = int(max(int(np.max(sl)) for sl in sls)) + 1
    mat = sparse_init(sls, nv)
    mat.data = np.zeros(mat.nnz, dtype=dtype)
    if cache_fn is not None:
        d = os.path.dirname(cache_fn)
        if d:
            os.makedirs(d, exist_ok=True)
        sp.save_npz(cache_fn, mat)
    return mat


def _csc_call(fn, X, Y, sls, mat, sls_Y, weights, *extra):
    """Runs a CSC searchlight entry point (TMat X, Y; TSearchlights; TSparseCSC;
    per-searchlight weights) and copies the accumulated data back into mat."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    hx = _SlHolder(sls)
    hy = _SlHolder(sls_Y) if sls_Y is not None else None
    data = np.ascontiguousarray(mat.data, dtype=np.float64)
    indices = np.ascontiguousarray(mat.indices, dtype=np.int32)
    indptr = np.ascontiguousarray(mat.indptr, dtype=np.int32)
    csc = TSparseCSC(_ptr(data, ctypes.c_double), _ptr(indices, ctypes.c_int32), _ptr(indptr, ctypes.c_int32),
                     mat.shape[0], mat.shape[1], data.size)
    tx = TMat(_ptr(X, ctypes.c_double), X.shape[0], X.shape[1])
    ty = TMat(_ptr(Y, ctypes.c_double), Y.shape[0], Y.shape[1])
    wp = None
    if weights is not None:
        warr = [np.ascontiguousarray(w, dtype=np.float64) for w in weights]
        wp = (ctypes.POINTER(ctypes.c_double) * len(warr))(*[_ptr(w, ctypes.c_double) for w in warr])
    rc = getattr(lib(), fn)(ctypes.byref(tx), ctypes.byref(ty), ctypes.byref(hx.st),
                            ctypes.byref(hy.st) if hy else None, ctypes.byref(csc), wp, *extra)
    _check(fn, rc)
    mat.data[:] = data
    return mat


def searchlight_procrustes_csc(X, Y, sls, mat, sls_Y=None, weights=None, reflection=True, scaling=False):
    """ha_searchlight_procrustes (CSC API) with backend HIP; mat is accumulated in place.

    mat: scipy.sparse.csc_matrix with sorted indices (e.g. from sparse_init).
    """
    return _csc_call("ha_searchlight_procrustes", X, Y, sls, mat, sls_Y, weights, bool(reflection), bool(scaling),
                     BACKENDS["hip"])


def searchlight_ridge_csc(X, Y, sls, mat, sls_Y=None, weights=None, alpha=1e3):
    """ha_searchlight_ridge (csrc/ha_searchlight.c:1378-1429) on the GPU: FP64
    ridge betas per searchlight, accumulated into mat in place (row weights,
    as the C scatter)."""
    return _csc_call("ha_searchlight_ridge", X, Y, sls, mat, sls_Y, weights, float(alpha))


def searchlight_ridge(X, Y, sls, sls_Y=None, mat0=None, alpha=1e3, weights=None):
    """searchlight_ridge of src/hyperalignment/searchlight.py:176-212 on the GPU.

    mat0: scipy CSC matrix whose pattern holds every searchlight pair (e.g.
    from initialize_sparse_matrix); its copy is accumulated and returned. With
    mat0 None the pattern is built for the searchlights and a dense array is
    returned, as the reference returns a dense array then. weights multiply the
    rows of each local betas matrix (the C library's scatter,
    csrc/ha_sparse.c:128-142); the Python reference multiplies columns
    (searchlight.py:125).
    """
    import scipy.sparse as sp
    if mat0 is None:
        if X.shape[1] != Y.shape[1] or sls_Y is not None:
            raise ValueError("mat0 is required with sls_Y or when X and Y have different numbers of vertices "
                             "(its pattern must hold every X x Y pair)")
        mat = initialize_sparse_matrix(sls, nv=X.shape[1])
        searchlight_ridge_csc(X, Y, sls, mat, None, weights, alpha)
        return mat.toarray()
    mat = sp.csc_matrix(mat0, dtype=np.float64, copy=True)
    mat.sort_indices()
    return searchlight_ridge_csc(X, Y, sls, mat, sls_Y, weights, alpha)


def mesh_searchlights_flat(coords, faces, radius, centers=None, geodesic=True):
    """Searchlights of `radius` around each centre (every vertex by default), as
    the flat arrays the searchlight entry points take: (indices int32, offsets
    int32, dists float64), each searchlight sorted by vertex id. geodesic:
    distances across the faces (ha_mesh_searchlights_geodesic, fast marching);
    else along the mesh edges (ha_mesh_searchlights, Dijkstra)."""
    V = np.ascontiguousarray(coords, dtype=np.float64)
    F = np.ascontiguousarray(faces, dtype=np.int32)
    if V.ndim != 2 or V.shape[1] != 3 or F.ndim != 2 or F.shape[1] != 3:
        raise ValueError("coords must be (nv, 3) and faces (nf, 3)")
    C = None if centers is None else np.ascontiguousarray(centers, dtype=np.int32)
    pi, po = ctypes.POINTER(ctypes.c_int32)(), ctypes.POINTER(ctypes.c_int32)()
    pd = ctypes.POINTER(ctypes.c_double)()
    nc = V.shape[0] if C is None else len(C)
    L = lib()
    name = "ha_mesh_searchlights_geodesic" if geodesic else "ha_mesh_searchlights"
    _check(name, getattr(L, name)(
        _ptr(V, ctypes.c_double), V.shape[0], _ptr(F, ctypes.c_int32), F.shape[0], float(radius),
        _ptr(C, ctypes.c_int32), 0 if C is None else len(C), ctypes.byref(pi), ctypes.byref(po), ctypes.byref(pd)))
    try:
        off = np.ctypeslib.as_array(po, shape=(nc + 1,)).copy()
        total = int(off[-1])
        idx = np.ctypeslib.as_array(pi, shape=(max(total, 1),))[:total].copy()
        dists = np.ctypeslib.as_array(pd, shape=(max(total, 1),))[:total].copy()
    finally:
        for ptr in (pi, po, pd):
            L.ha_mesh_free(ctypes.cast(ptr, ctypes.c_void_p))
    return idx, off, dists


def mesh_searchlights(coords, faces, radius, centers=None, return_dists=False, geodesic=True):
    """Searchlights on a triangle mesh in the form neuroboros' nb.sls returns
    them (benchmark_neuroboros.py:153): a list of vertex-id arrays, and with
    return_dists the matching distance arrays. Distances are geodesic across
    the faces (fast marching), or along the mesh edges with geodesic=False
    (Dijkstra); a searchlight holds the vertices closer than radius."""
    idx, off, d = mesh_searchlights_flat(coords, faces, radius, centers, geodesic)
    sls = [idx[off[i]:off[i + 1]] for i in range(len(off) - 1)]
    if not return_dists:
        return sls
    return sls, [d[off[i]:off[i + 1]] for i in range(len(off) - 1)]


def read_surface(path):
    """(coords (nv, 3) float64, faces (nf, 3) int32) of a FreeSurfer triangle
    surface file (lh.white, ...) or a GIFTI surface (.gii: the POINTSET and
    TRIANGLE data arrays; ASCII, Base64Binary or GZipBase64Binary encodings)."""
    with open(path, "rb") as fh:
        head = fh.read(3)
    if head == b"\xff\xff\xfe":
        return _read_freesurfer(path)
    return _read_gifti(path)


def _read_freesurfer(path):
    with open(path, "rb") as fh:
        fh.read(3)
        fh.readline()  # creator line
        fh.readline()  # blank line
        nv, nf = np.frombuffer(fh.read(8), dtype=">i4")
        coords = np.frombuffer(fh.read(12 * int(nv)), dtype=">f4").reshape(int(nv), 3)
        faces = np.frombuffer(fh.read(12 * int(nf)), dtype=">i4").reshape(int(nf), 3)
    return coords.astype(np.float64), faces.astype(np.int32)


def _read_gifti(path):
    import base64
    import xml.etree.ElementTree as ET
    import zlib
    dtypes = {"NIFTI_TYPE_FLOAT32": np.float32, "NIFTI_TYPE_FLOAT64": np.float64, "NIFTI_TYPE_INT32": np.int32,
              "NIFTI_TYPE_UINT8": np.uint8}
    out = {}
    for da in ET.parse(path).getroot().iter("DataArray"):
        intent = da.get("Intent")
        if intent not in ("NIFTI_INTENT_POINTSET", "NIFTI_INTENT_TRIANGLE"):
            continue
        dims = [int(da.get(f"Dim{i}")) for i in range(int(da.get("Dimensionality")))]
        dt = np.dtype(dtypes[da.get("DataType")])
        if da.get("Endian", "LittleEndian") == "BigEndian":
            dt = dt.newbyteorder(">")
        text = da.find("Data").text or ""
        enc = da.get("Encoding", "ASCII")
        if enc == "ASCII":
            a = np.array(text.split(), dtype=dt)
        else:
            raw = base64.b64decode("".join(text.split()))
            if enc == "GZipBase64Binary":
                raw = zlib.decompress(raw, 15 + 32)  # zlib stream (the standard's) or gzip framing
            a = np.frombuffer(raw, dtype=dt)
        order = "F" if da.get("ArrayIndexingOrder") == "ColumnMajorOrder" else "C"
        out[intent] = a.reshape(dims, order=order)
    return (out["NIFTI_INTENT_POINTSET"].astype(np.float64), out["NIFTI_INTENT_TRIANGLE"].astype(np.int32))


def ridge(X, Y, alpha):
    """ridge of src/hyperalignment/ridge.py:7-12 / csrc/ha_ridge.c:13-54 on the
    GPU (FP64 normal equations, the same betas for alpha > 0). Y may be 1-D
    (ha_ridge_vec, csrc/ha_ridge.c:56-95)."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    tx = TMat(_ptr(X, ctypes.c_double), X.shape[0], X.shape[1])
    if Y.ndim == 1:
        b = np.zeros(X.shape[1], dtype=np.float64)
        _check("ha_ridge_vec", lib().ha_ridge_vec(ctypes.byref(tx), _ptr(Y, ctypes.c_double), float(alpha),
                                                  _ptr(b, ctypes.c_double)))
        return b
    b = np.zeros((X.shape[1], Y.shape[1]), dtype=np.float64)
    ty = TMat(_ptr(Y, ctypes.c_double), Y.shape[0], Y.shape[1])
    tb = TMat(_ptr(b, ctypes.c_double), b.shape[0], b.shape[1])
    _check("ha_ridge", lib().ha_ridge(ctypes.byref(tx), ctypes.byref(ty), float(alpha), ctypes.byref(tb)))
    return b


def sparse_scatter_add(mat, rows, cols, local_T, weights=None):
    """ha_sparse_scatter_add on the GPU (one local transform into a CSC matrix)."""
    data = np.ascontiguousarray(mat.data, dtype=np.float64)
    indices = np.ascontiguousarray(mat.indices, dtype=np.int32)
    indptr = np.ascontiguousarray(mat.indptr, dtype=np.int32)
    csc = TSparseCSC(_ptr(data, ctypes.c_double), _ptr(indices, ctypes.c_int32), _ptr(indptr, ctypes.c_int32),
                     mat.shape[0], mat.shape[1], data.size)
    r = np.ascontiguousarray(rows, dtype=np.int32)
    c = np.ascontiguousarray(cols, dtype=np.int32)
    T = np.ascontiguousarray(local_T, dtype=np.float64)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    lib().ha_sparse_scatter_add(ctypes.byref(csc), _ptr(r, ctypes.c_int32), _ptr(c, ctypes.c_int32), len(r),
                                len(c), _ptr(T, ctypes.c_double), _ptr(w, ctypes.c_double))
    mat.data[:] = data
    return mat


def procrustes(X, Y, reflection=True, scaling=False):
    """ha_procrustes (one local problem, FP64 Jacobi on the GPU)."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    T = np.zeros((X.shape[1], X.shape[1]), dtype=np.float64)
    tx = TMat(_ptr(X, ctypes.c_double), X.shape[0], X.shape[1])
    ty = TMat(_ptr(Y, ctypes.c_double), Y.shape[0], Y.shape[1])
    tt = TMat(_ptr(T, ctypes.c_double), T.shape[0], T.shape[1])
    _check("ha_procrustes", lib().ha_procrustes(ctypes.byref(tx), ctypes.byref(ty), ctypes.byref(tt),
                                                bool(reflection), bool(scaling)))
    return T


def polar_newton(A, reflection=True, scaling=False, X=None):
    """ha_polar_newton (csrc/ha_procrustes.c:271-366) on the GPU: the orthogonal
    polar factor of the square float32 matrix A, with the reference's
    reflection rule (det < 0: last column negated) and scaling by
    trace(T^T A) / (sum var(X) * M) when X (M x N) is given."""
    A = np.ascontiguousarray(A, dtype=np.float32)
    n = A.shape[0]
    if A.shape != (n, n):
        raise ValueError("A must be square")
    T = np.zeros((n, n), dtype=np.float32)
    Xf = None if X is None else np.ascontiguousarray(X, dtype=np.float32)
    m = 0 if Xf is None else Xf.shape[0]
    _check("ha_polar_newton", lib().ha_polar_newton(_ptr(A, ctypes.c_float), _ptr(T, ctypes.c_float), n,
                                                    bool(reflection), bool(scaling), _ptr(Xf, ctypes.c_float), m))
    return T


class Plan:
    """Device-resident searchlight pass (ha_hip_plan_*): searchlights and data in HBM."""

    def __init__(self, nt, nv, idx, offsets, dists, radius):
        self._idx = np.ascontiguousarray(idx, dtype=np.int32)
        self._off = np.ascontiguousarray(offsets, dtype=np.int32)
        self._d = None if dists is None else np.ascontiguousarray(dists, dtype=np.float64)
        self.nt, self.nv = int(nt), int(nv)
        self.count = len(self._off) - 1
        h = ctypes.c_void_p()
        rc = lib().ha_hip_plan_create(self.nt, self.nv, _ptr(self._idx, ctypes.c_int32),
                                      _ptr(self._off, ctypes.c_int32), _ptr(self._d, ctypes.c_double),
                                      self.count, float(radius), ctypes.byref(h))
        _check("ha_hip_plan_create", rc)
        self.h = h

    def load(self, X, Y):
        X = np.asfortranarray(X, dtype=np.float64)
        Y = np.asfortranarray(Y, dtype=np.float64)
        _check("ha_hip_plan_load", lib().ha_hip_plan_load(self.h, _ptr(X, ctypes.c_double),
                                                          _ptr(Y, ctypes.c_double), True))

    def execute(self, reflection=True, scaling=False):
        _check("ha_hip_plan_execute", lib().ha_hip_plan_execute(self.h, bool(reflection), bool(scaling)))

    def ridge(self, alpha):
        """Searchlight ridge pass into the plan's CSC data (ha_hip_plan_ridge)."""
        _check("ha_hip_plan_ridge", lib().ha_hip_plan_ridge(self.h, float(alpha)))

    def nnz(self):
        n = ctypes.c_int64()
        _check("ha_hip_plan_nnz", lib().ha_hip_plan_nnz(self.h, ctypes.byref(n)))
        return n.value

    def csc(self, with_data=True):
        nnz = self.nnz()
        indptr = np.zeros(self.nv + 1, dtype=np.int32)
        indices = np.zeros(max(nnz, 1), dtype=np.int32)
        data = np.zeros(max(nnz, 1), dtype=np.float64) if with_data else None
        _check("ha_hip_plan_download", lib().ha_hip_plan_download(self.h, _ptr(indptr, ctypes.c_int32),
                                                                  _ptr(indices, ctypes.c_int32),
                                                                  _ptr(data, ctypes.c_double)))
        return indptr, indices[:nnz], (data[:nnz] if with_data else None)

    def local(self, which):
        which = np.ascontiguousarray(which, dtype=np.int32)
        sizes = np.diff(self._off)[which].astype(np.int64)
        out = np.zeros(int((sizes ** 2).sum()), dtype=np.float32)
        _check("ha_hip_plan_download_local", lib().ha_hip_plan_download_local(
            self.h, _ptr(which, ctypes.c_int32), len(which), _ptr(out, ctypes.c_float)))
        res, o = [], 0
        for k in sizes:
            res.append(out[o:o + k * k].reshape(k, k))
            o += k * k
        return res

    def spectra(self, which):
        """(features (n, 12), [sigma_max, sigma_min] (n, 2)) from the FP32 Jacobi path."""
        which = np.ascontiguousarray(which, dtype=np.int32)
        f = np.zeros((len(which), MLP_IN), dtype=np.float32)
        s = np.zeros((len(which), 2), dtype=np.float32)
        _check("ha_hip_plan_spectra", lib().ha_hip_plan_spectra(self.h, _ptr(which, ctypes.c_int32), len(which),
                                                                _ptr(f, ctypes.c_float), _ptr(s, ctypes.c_float)))
        return f, s

    def steps(self):
        """(steps, fallback flags) per searchlight of the last pass."""
        st = np.zeros(self.count, dtype=np.int32)
        fl = np.zeros(self.count, dtype=np.int32)
        _check("ha_hip_plan_steps", lib().ha_hip_plan_steps(self.h, _ptr(st, ctypes.c_int32),
                                                            _ptr(fl, ctypes.c_int32)))
        return st, fl

    def predictions(self):
        p = np.zeros((self.count, 2), dtype=np.float32)
        _check("ha_hip_plan_predictions", lib().ha_hip_plan_predictions(self.h, _ptr(p, ctypes.c_float)))
        return p

    def evaluate(self, X, Y=None, want_yhat=True):
        """Yhat = X @ T with the plan's CSC transform and, when Y is given, the
        vertex-wise correlation of Yhat with Y (benchmark_neuroboros.py:204-211,229-232).
        Returns (Yhat or None, r or None)."""
        X = np.asfortranarray(X, dtype=np.float64)
        nt = X.shape[0]
        Yf = None if Y is None else np.asfortranarray(Y, dtype=np.float64)
        yh = np.zeros((nt, self.nv), dtype=np.float64, order="F") if want_yhat else None
        r = None if Y is None else np.zeros(self.nv, dtype=np.float64)
        _check("ha_hip_plan_evaluate", lib().ha_hip_plan_evaluate(
            self.h, _ptr(X, ctypes.c_double), _ptr(Yf, ctypes.c_double), nt, True,
            _ptr(yh, ctypes.c_double), _ptr(r, ctypes.c_double)))
        return (None if yh is None else np.ascontiguousarray(yh)), r

    def tiles(self):
        """Iteration tile per kp bucket (kp <= 64, 128, 256, 384, 512, larger)."""
        t = np.zeros(6, dtype=np.int32)
        _check("ha_hip_plan_tiles", lib().ha_hip_plan_tiles(self.h, _ptr(t, ctypes.c_int32)))
        return t

    def tuning(self):
        """The autotuner's launch per kp bucket (kp <= 64, 128, 256, 384, 512,
        larger) and the sample times behind it: a list of dicts with the tile,
        whether 128 x 128 tiles run in teams, the searchlights of the bucket, the
        sample timed and its iteration time per candidate (ms; None where not
        timed)."""
        t = HaHipTuning()
        _check("ha_hip_plan_tuning", lib().ha_hip_plan_tuning(self.h, ctypes.byref(t)))
        edges = ["kp<=64", "kp<=128", "kp<=256", "kp<=384", "kp<=512", "kp>512"]
        out = []
        for b in range(6):
            timed = t.sample[b] > 0
            out.append({"bucket": edges[b], "searchlights": int(t.searchlights[b]), "tile": int(t.tile[b]),
                        "teams": bool(t.team[b]), "sample": int(t.sample[b]),
                        "ms": ({"64": float(t.ms[b][0]), "128": float(t.ms[b][1]),
                                "128_teams": float(t.ms[b][2]) or None} if timed else None)})
        return out

    def det_signs(self, lu_kernel=True):
        """1 where the last pass's local transform has det Q_s < 0, else 0, per
        searchlight: from the LU-sign kernel (lu_kernel) or from the LU inside
        the reflection-fix kernel."""
        out = np.zeros(self.count, dtype=np.int32)
        _check("ha_hip_plan_det_signs", lib().ha_hip_plan_det_signs(self.h, 1 if lu_kernel else 0,
                                                                    _ptr(out, ctypes.c_int32)))
        return out

    def orthogonality(self, exact=True):
        """||Q_s^T Q_s - I||_F of every local transform in FP64 on the GPU (exact:
        every Gram entry; else the four-probe estimate the pass uses)."""
        out = np.zeros(self.count, dtype=np.float64)
        _check("ha_hip_plan_orthogonality", lib().ha_hip_plan_orthogonality(self.h, 1 if exact else 0,
                                                                            _ptr(out, ctypes.c_double)))
        return out

    def stats(self):
        s = HaHipStats()
        _check("ha_hip_plan_stats", lib().ha_hip_plan_stats(self.h, ctypes.byref(s)))
        return s.as_dict()

    def close(self):
        if getattr(self, "h", None):
            lib().ha_hip_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
