"""Why the template kernel forms A and the polar factor in FP64 (host emulation, no GPU).

The template chain (ha_template.c:14-114) runs 47 dependent Procrustes solves per
searchlight on the 16-subject workload (order-5 mesh, n = 400, 20 mm). Most of its
cross-covariances have sigma_min / sigma_max around 2.5e-4, but some reach 1e-8: in
searchlight 2435 of that data one
subject's A has 1.8e-8. An FP32 polar iteration cannot resolve that smallest
singular pair, and the local template then leaves the FP64 reference by O(1).

The emulation runs the chain with the kernel's arithmetic choices: the subjects
and the aligned data stored in FP32, A = D_s^T R and the scheduled polar
iteration (the library's own schedule table, ha_polar_schedule; stop at
||G - I||_F < tol, then one quintic Newton-Schulz step) either in FP32 (the
previous kernel) or in FP64 (the current one), against the oracle's LAPACK
template.
"""

import ctypes

import numpy as np
import pytest

import hyperalignment_amd as ha
from hyperalignment_amd import synth
from oracle import ha_oracle as O

BAR = 1e-4   # max-abs bar of the template tests, relative to the largest magnitude


def schedule(ell, steps=40):
    fn = ha.lib().ha_polar_schedule
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int)]
    out = np.zeros(steps * 3, dtype=np.float32)
    gl = ctypes.c_int()
    n = fn(ell, steps, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ctypes.byref(gl))
    return out.reshape(steps, 3)[:n].astype(np.float64)


def polar(A, dt, co, tol):
    """The scheduled iteration X <- X (a I + b G + c G^2) from A / ||A||_F in dtype dt."""
    k = A.shape[0]
    X = (A / np.linalg.norm(A)).astype(dt)
    I = np.eye(k, dtype=dt)
    for t in range(80):
        G = X.T @ X
        finish = np.linalg.norm(G.astype(np.float64) - np.eye(k)) < tol
        a, b, c = (1.875, -1.25, 0.375) if finish or t >= len(co) else co[t]
        X = (X @ (dt(a) * I + G @ (dt(b) * I + dt(c) * G))).astype(dt)
        if finish:
            return X.astype(np.float64)
    raise AssertionError("no convergence")


def zscore(X, dt):
    X = X.astype(np.float64)
    return ((X - X.mean(0)) / X.std(0)).astype(dt)


def chain(loc, fp64, co, level2=2):
    """The template of ha_template.c:14-114 (z-scored, level2 passes) with FP32
    storage; A and the iteration in FP64 when fp64, else in FP32."""
    f32 = np.float32
    D = [d.astype(f32) for d in loc]
    ns = len(D)

    def solve(Ds, R):
        if fp64:
            A = Ds.astype(np.float64).T @ R.astype(np.float64)
            Q = polar(A, np.float64, co, 1e-3)
        else:
            A = (Ds.T @ R).astype(f32)
            Q = polar(A.astype(np.float64), f32, co, 1e-2)
        return (Ds @ Q.astype(f32)).astype(f32)

    common = D[0].copy()
    aligned = [D[0].copy()]
    for s in range(1, ns):
        a = zscore(solve(D[s], common), f32)
        aligned.append(a)
        common = zscore((common + a) * f32(0.5), f32)
    for _ in range(level2):
        total = sum(aligned[1:], aligned[0].copy())
        for s in range(ns):
            ref = zscore((total - aligned[s]) * f32(1.0 / (ns - 1)), f32)
            aligned[s] = solve(D[s], ref)
    return zscore(sum(aligned[1:], aligned[0].copy()), np.float64)


@pytest.fixture(scope="module")
def case():
    dms, sls = synth.template_subjects(n_subj=16, n=400, order=5, radius=20.0, seed=0)
    cols = sls.lists()[2435]
    loc = [d[:, cols] for d in dms]
    return loc, O.template_procrustes(loc, level2_iter=2), schedule(1e-7)


def test_case_is_ill_conditioned(case):
    """Along the oracle's first pass (each subject against the running common
    space) one cross-covariance has sigma_min / sigma_max below 1e-7."""
    loc, _, _ = case
    common = loc[0].copy()
    worst = 1.0
    for d in loc[1:]:
        s = np.linalg.svd(d.T @ common, compute_uv=False)
        worst = min(worst, s[-1] / s[0])
        a = O.zscore_columns(d @ O.procrustes(d, common))
        common = O.zscore_columns((common + a) * 0.5)
    print(f"smallest sigma_min / sigma_max of the first pass: {worst:.1e}")
    assert worst < 1e-7


def test_fp32_iteration_loses_the_template(case):
    loc, ref, co = case
    err = np.abs(chain(loc, False, co) - ref).max() / np.abs(ref).max()
    print(f"FP32 A and iteration: {err:.2e} of the largest magnitude")
    assert err > 100 * BAR


def test_fp64_iteration_keeps_the_bar(case):
    loc, ref, co = case
    err = np.abs(chain(loc, True, co) - ref).max() / np.abs(ref).max()
    print(f"FP64 A and iteration: {err:.2e} of the largest magnitude")
    assert err < BAR
