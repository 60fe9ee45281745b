"""The C/OpenMP restatement of the reference's CPU FP64 searchlight Procrustes
(oracle/cpu_ref.c; driver ha_searchlight.c:950-1106, local solve
ha_searchlight.c:578-659), for the bench's CPU baseline.

Test infrastructure: only tests/ and bench.py's cpu_baseline use it. BLAS and
LAPACK are scipy's (LP64 interface, handed to the C code as function
pointers), one BLAS thread per OpenMP thread.
"""

import ctypes
import os

import numpy as np

_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_DIR, "libha_cpu_ref.so")
_lib = None


def _fn_ptr(mod, name):
    cap = mod.__pyx_capi__[name]
    get = ctypes.pythonapi.PyCapsule_GetPointer
    get.restype = ctypes.c_void_p
    get.argtypes = [ctypes.py_object, ctypes.c_char_p]
    get_name = ctypes.pythonapi.PyCapsule_GetName
    get_name.restype = ctypes.c_char_p
    get_name.argtypes = [ctypes.py_object]
    return get(cap, get_name(cap))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} not built (make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        L.cpu_ref_procrustes.restype = ctypes.c_int
        P = ctypes.c_void_p
        L.cpu_ref_procrustes.argtypes = [P, P, P, P, P, P, P, ctypes.c_int, P, P, P, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, P, ctypes.c_int, P]
        _lib = L
    return _lib


def core_cpus():
    """One CPU of each physical core this process may run on (from the
    affinity mask and /sys topology; CPUs without topology count as cores)."""
    import glob
    avail = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else list(range(os.cpu_count() or 1))
    seen, out = set(), []
    for c in avail:
        key = None
        try:
            pkg = open(f"/sys/devices/system/cpu/cpu{c}/topology/physical_package_id").read().strip()
            core = open(f"/sys/devices/system/cpu/cpu{c}/topology/core_id").read().strip()
            key = (pkg, core)
        except OSError:
            key = ("cpu", c)
        if key not in seen:
            seen.add(key)
            out.append(c)
    return out


def procrustes_searchlights(X, Y, indices, offsets, which, threads, reflection=True, cpus=None):
    """Local transforms (k x k each) of the searchlights `which`, computed by
    the C restatement on `threads` OpenMP threads, thread t bound to cpus[t]
    (None: unbound); X, Y: nt x nv FP64."""
    import scipy.linalg.cython_blas as cb
    import scipy.linalg.cython_lapack as cl
    Xf = np.asfortranarray(X, dtype=np.float64)
    Yf = np.asfortranarray(Y, dtype=np.float64)
    idx = np.ascontiguousarray(indices, dtype=np.int32)
    off = np.ascontiguousarray(offsets, dtype=np.int32)
    wh = np.ascontiguousarray(which, dtype=np.int32)
    sizes = (off[wh + 1] - off[wh]).astype(np.int64)
    out = np.zeros(int((sizes ** 2).sum()), dtype=np.float64)
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    cp = np.ascontiguousarray(cpus if cpus is not None else [0], dtype=np.int32)
    rc = lib().cpu_ref_procrustes(
        _fn_ptr(cb, "dgemm"), _fn_ptr(cl, "dgesdd"), _fn_ptr(cl, "dgesvd"), _fn_ptr(cl, "dgetrf"),
        _fn_ptr(cb, "dger"), ptr(Xf), ptr(Yf), int(X.shape[0]), ptr(idx), ptr(off), ptr(wh), len(wh),
        1 if reflection else 0, int(threads), ptr(cp), len(cpus) if cpus is not None else 0, ptr(out))
    if rc != 0:
        raise RuntimeError(f"cpu_ref_procrustes failed ({rc})")
    res, o = [], 0
    for k in sizes:
        res.append(out[o:o + k * k].reshape(k, k).T)  # column-major k x k
        o += k * k
    return res
