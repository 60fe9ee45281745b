"""The polar coefficient schedules built by the library (host code, no GPU)."""

import ctypes

import numpy as np
import pytest

import hyperalignment_amd as ha


def schedule(ell, steps=28):
    L = ha.lib()
    fn = L.ha_polar_schedule
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int)]
    out = np.zeros(steps * 3, dtype=np.float32)
    gl = ctypes.c_int()
    n = fn(ell, steps, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ctypes.byref(gl))
    assert n == steps
    return out.reshape(steps, 3).astype(np.float64), gl.value


@pytest.mark.parametrize("ell", [10 ** (-b / 4) for b in range(1, 37, 5)])
def test_schedule_converges_and_never_crushes(ell):
    c, growth = schedule(ell)
    x = np.concatenate([np.geomspace(ell, 1.0, 400), np.linspace(0.5, 1.01, 200)])
    lo = np.inf
    for t in range(len(c)):
        a, b, cc = c[t]
        x = a * x + b * x ** 3 + cc * x ** 5
        if t >= 1:
            lo = min(lo, x.min())
        if t >= growth + 4:
            break
    assert np.abs(x - 1).max() < 1e-6
    # after the first step no singular value falls below the floor image
    assert lo > min(0.1, ell)
    # every value stays positive (no sign flip) and bounded
    assert x.min() > 0 and x.max() < 1.02


def test_growth_length_monotone():
    lens = [schedule(10 ** (-b / 4))[1] for b in range(1, 37)]
    assert all(a <= b for a, b in zip(lens, lens[1:]))
    assert lens[-1] <= 16
