"""ha_procrustes_noalloc: ha_procrustes on device buffers with no allocation
(its work lives in the library's static scratch array). Checked against the
allocating entry point ha_procrustes on the same inputs, through the C-ABI:
two calls in a row on one stream, the second smaller than the first, for each
flag combination; both run the same FP64 Jacobi kernel, so the bar is
max-abs 1e-12 (bit-identical in practice). Sizes beyond the stated maximum are
rejected with kHaErrorArg. The device buffers here come from hipMalloc of the
HIP runtime the library itself is linked against (the test may allocate; the
entry point must not).
"""

import ctypes
import os

import numpy as np
import pytest

import hyperalignment_amd as ha

pytestmark = pytest.mark.gpu

MAX_N = 512
MAX_MN = 1 << 20


def _hip():
    for path in ("/opt/rocm/lib/libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so"):
        if os.path.exists(path):
            h = ctypes.CDLL(path)
            break
    else:
        pytest.skip("HIP runtime library not found")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    h.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    h.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    for fn in ("hipMalloc", "hipFree", "hipMemcpy", "hipStreamSynchronize", "hipStreamCreate",
               "hipStreamDestroy", "hipDeviceSynchronize"):
        getattr(h, fn).restype = ctypes.c_int
    return h


H2D, D2H = 1, 2


class Dev:
    """A device buffer of float64 from hipMalloc (freed on close)."""

    def __init__(self, hip, n):
        self.hip = hip
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), max(8, 8 * n)) == 0
        self.n = n

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size <= self.n
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, 8 * a.size, H2D) == 0

    def get(self, n, shape):
        out = np.zeros(n, dtype=np.float64)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, 8 * n, D2H) == 0
        return out.reshape(shape)

    def close(self):
        self.hip.hipFree(self.p)


@pytest.fixture(scope="module")
def env():
    ha.init()
    hip = _hip()
    fn = ha.lib().ha_procrustes_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p,
                   ctypes.c_bool, ctypes.c_bool, ctypes.c_void_p]
    stream = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    bufs = [Dev(hip, MAX_MN), Dev(hip, MAX_MN), Dev(hip, MAX_N * MAX_N)]
    yield hip, fn, stream, bufs
    for b in bufs:
        b.close()
    hip.hipStreamDestroy(stream)


def _case(seed, M, N):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((M, N))
    R, _ = np.linalg.qr(rng.standard_normal((N, N)))
    Y = X @ R + 0.5 * rng.standard_normal((M, N))
    return X, Y


@pytest.mark.parametrize("reflection,scaling", [(True, False), (False, False), (True, True), (False, True)])
def test_noalloc_matches_ha_procrustes(env, reflection, scaling):
    hip, fn, stream, (dx, dy, dt) = env
    for seed, (M, N) in enumerate([(700, 300), (90, 41)]):   # the second call smaller than the first
        X, Y = _case(seed + (0 if reflection else 10) + (0 if not scaling else 20), M, N)
        dx.put(X)
        dy.put(Y)
        rc = fn(dx.p, dy.p, M, N, dt.p, reflection, scaling, stream)
        assert rc == 0
        assert hip.hipStreamSynchronize(stream) == 0
        got = dt.get(N * N, (N, N))
        ref = ha.procrustes(X, Y, reflection=reflection, scaling=scaling)
        err = np.abs(got - ref).max()
        print(f"[noalloc] M {M} N {N} reflection {reflection} scaling {scaling}: max |diff| {err:.3e}",
              flush=True)
        assert err <= 1e-12
        if not reflection:
            assert np.linalg.det(got) > 0


def test_noalloc_rejects_sizes_beyond_the_maximum(env):
    hip, fn, stream, (dx, dy, dt) = env
    assert fn(dx.p, dy.p, 4, MAX_N + 1, dt.p, True, False, stream) == -3
    assert fn(dx.p, dy.p, MAX_MN // 100 + 1, 100, dt.p, True, False, stream) == -3
    assert fn(None, dy.p, 4, 4, dt.p, True, False, stream) == -3
