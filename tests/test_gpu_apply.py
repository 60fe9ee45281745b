"""Applying the searchlight transform to new data and scoring it, on the GPU
(ha_hip_plan_evaluate) against the oracle. FP64 throughout: Yhat bit-exact
against the in-order accumulation on the small mesh, 1e-12 relative against
scipy on the larger one; correlations within 1e-12."""

import numpy as np
import pytest
import scipy.sparse as sp

import hyperalignment_amd as ha
from hyperalignment_amd import synth
from oracle import ha_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def hip():
    ha.init()
    yield
    ha.set_options(method="auto")


def _plan(name):
    w = synth.make_workload(name)
    s = w.sls
    plan = ha.Plan(w.X.shape[0], s.nv, s.indices, s.offsets, s.dists, s.radius)
    plan.load(w.X, w.Y)
    ha.set_options(method="polar" if name != "c0" else "jacobi64")
    plan.execute()
    return w, plan


def test_apply_bitexact_c0():
    w, plan = _plan("c0")
    try:
        indptr, indices, data = plan.csc()
        rng = np.random.default_rng(1)
        Xt = rng.standard_normal((37, w.sls.nv))
        Yhat, _ = plan.evaluate(Xt)
        ref = O.apply_transform(Xt, indptr, indices, data, w.sls.nv)
        assert np.array_equal(Yhat.view(np.uint64), ref.view(np.uint64))
    finally:
        plan.close()


def test_evaluate_c1():
    w, plan = _plan("c1")
    try:
        indptr, indices, data = plan.csc()
        nv = w.sls.nv
        W = sp.csc_matrix((data, indices, indptr), shape=(nv, nv))
        rng = np.random.default_rng(2)
        Xt = rng.standard_normal((120, nv))
        Yt = np.asarray(Xt @ W) + 0.5 * rng.standard_normal((120, nv))
        Yt[:, 5] = 3.0  # a constant column scores NaN
        Yhat, r = plan.evaluate(Xt, Yt)
        ref = np.asarray(Xt @ W)
        assert np.abs(Yhat - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
        rref = O.vertex_correlation(ref, Yt)
        ok = ~np.isnan(rref)
        assert np.isnan(r[5]) and np.array_equal(np.isnan(r), ~ok)
        assert np.abs(r[ok] - rref[ok]).max() < 1e-12
        # the transform carries signal: mean correlation well above zero
        assert np.nanmean(r) > 0.5
    finally:
        plan.close()
