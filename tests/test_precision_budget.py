"""Why every product of the polar iteration stays bf16x3 (host emulation, no GPU).

The iteration X <- X (a I + b G + c G^2), G = X^T X, runs with the library's
own coefficient schedule (ha_polar_schedule) and stopping rule (||G - I||_F <
1e-2, then one quintic Newton-Schulz step), emulated in numpy with the
arithmetic of the kernels:

* bf16x3: every stored matrix is an FP32 value split into bf16 hi and lo
  planes, every product is hi*hi + lo*hi + hi*lo accumulated in FP32 (the
  scheme of k_gemm.h);
* bf16: every stored matrix is rounded to one bf16 plane and every product is
  a single bf16 MFMA with FP32 accumulation;
* bf16 middle: single bf16 for the first half of the steps, bf16x3 after.

On cross-covariance-like matrices (singular values spread over 2.5 decades,
about the spread of the full-hemisphere searchlights) bf16x3 stays within the 1e-4 bar of the exact polar
factor, and any single-bf16 step leaves an error of 1e-3 or more that the later
bf16x3 steps do not remove: a rounding of the iterate at step t moves its polar
factor by about u / (s_i + s_j) in the (i, j) singular directions, and the
iteration then converges to the polar factor of the rounded iterate.
"""

import ctypes

import numpy as np
import pytest

import hyperalignment_amd as ha

TOL = 1e-2      # the library's default ||G - I||_F stopping tolerance
BAR = 1e-4      # max |Q - Q_exact| of the parity tests


def schedule(ell, steps=28):
    fn = ha.lib().ha_polar_schedule
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int)]
    out = np.zeros(steps * 3, dtype=np.float32)
    gl = ctypes.c_int()
    assert fn(ell, steps, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ctypes.byref(gl)) == steps
    return out.reshape(steps, 3)


def bf16(x):
    """Round-to-nearest-even FP32 -> bf16, returned as FP32 (the kernels' f2bf)."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def product(A, B, mode):
    """A @ B with the arithmetic of one precision mode (FP32 accumulation)."""
    if mode == "bf16":
        return bf16(A) @ bf16(B)
    ah, bh = bf16(A), bf16(B)
    al, bl = bf16(A - ah), bf16(B - bh)
    return ah @ bh + al @ bh + ah @ bl


def store(M, mode):
    """What a stored iterate holds: bf16 hi + lo planes (bf16x3), or one bf16 plane."""
    M = M.astype(np.float32)
    if mode == "bf16":
        return bf16(M)
    h = bf16(M)
    return h + bf16(M - h)


def polar_emulated(A, ell, mode_of_step, steps=None):
    """The scheduled iteration from X0 = A / sigma_max; mode_of_step(t) picks the
    arithmetic of step t. It stops as the kernels do, or, with `steps` given,
    after exactly that many steps (the last one the finishing step), since
    single-bf16 iterates never reach the stopping tolerance. Returns (Q, steps)."""
    k = A.shape[0]
    co = schedule(ell)
    X = store(A / np.linalg.norm(A, 2), mode_of_step(0))
    I = np.eye(k, dtype=np.float32)
    for t in range(40 if steps is None else steps):
        m = mode_of_step(t)
        G = store(product(X.T, X, m), m)
        if steps is None:
            finish = np.linalg.norm(G.astype(np.float64) - np.eye(k)) < TOL
        else:
            finish = t == steps - 1
        a, b, c = (1.875, -1.25, 0.375) if finish or t >= len(co) else co[t]
        P = store(a * I + product(G, b * I + c * G, m), m)
        X = store(product(X, P, m), m)
        if finish:
            return X.astype(np.float64), t + 1
    raise AssertionError("no convergence")


def make_case(seed, k=64, decades=2.5):
    rng = np.random.default_rng(seed)
    U, _ = np.linalg.qr(rng.standard_normal((k, k)))
    V, _ = np.linalg.qr(rng.standard_normal((k, k)))
    s = np.geomspace(1.0, 10 ** -decades, k)
    return (U * s) @ V.T, U @ V.T, 10 ** -decades


SEEDS = [0, 1, 2]


@pytest.mark.parametrize("seed", SEEDS)
def test_bf16x3_iteration_keeps_the_bar(seed):
    A, Q, ell = make_case(seed)
    got, steps = polar_emulated(A, ell, lambda t: "bf16x3")
    err = np.abs(got - Q).max()
    print(f"bf16x3: max|Q - Q_exact| = {err:.2e} in {steps} steps")
    assert err < BAR


@pytest.mark.parametrize("seed", SEEDS)
def test_single_bf16_anywhere_breaks_the_bar(seed):
    A, Q, ell = make_case(seed)
    _, n = polar_emulated(A, ell, lambda t: "bf16x3")
    everywhere, _ = polar_emulated(A, ell, lambda t: "bf16", steps=n + 2)
    middle, _ = polar_emulated(A, ell, lambda t: "bf16" if t < n // 2 else "bf16x3", steps=n + 2)
    e_all, e_mid = np.abs(everywhere - Q).max(), np.abs(middle - Q).max()
    print(f"single bf16: everywhere {e_all:.2e}, first half then bf16x3 {e_mid:.2e}")
    assert e_all > 10 * BAR
    assert e_mid > 10 * BAR


def test_damage_is_not_rounding_of_the_output():
    """The single-bf16 error is not the last rounding: with the last four steps
    in bf16x3 (two more than the bf16x3 iteration needs) the error stays more
    than ten times the bar, because the iterate's polar factor itself moved."""
    A, Q, ell = make_case(7)
    _, n = polar_emulated(A, ell, lambda t: "bf16x3")
    got, _ = polar_emulated(A, ell, lambda t: "bf16" if t < n - 2 else "bf16x3", steps=n + 2)
    assert np.abs(got - Q).max() > 10 * BAR
