"""The sign of det Q that the reflection fix (isReflection = false,
csrc/ha_searchlight.c:612-632) turns on, from the LU-sign kernel (k_lu_sign:
panels in LDS, row interchanges as a permutation, trailing updates on FP32
MFMA) against the LU inside the reflection-fix kernel and against numpy's
FP64 determinant of the downloaded local transforms.

The bar is exact agreement: Q is orthogonal, |det Q| = 1, so an FP32 LU with
partial pivoting decides the sign with a margin of many orders of magnitude.
Workloads: the 10,242-vertex mesh (k up to 441, kp 448: two workgroups per CU),
the full hemisphere (all 40,962 searchlights), and the long-tail variant
(kp up to 544: the panel then takes one workgroup per CU).
"""

import numpy as np
import pytest

import hyperalignment_amd as ha
from hyperalignment_amd import learn, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def hip():
    ha.init()
    learn.install(learn.MODEL_PATH)
    ha.set_options(method="auto")
    yield
    ha.set_options(method="auto")


def _numpy_negative(plan, which):
    return np.array([np.linalg.det(q.astype(np.float64)) < 0 for q in plan.local(which)], dtype=np.int32)


@pytest.mark.parametrize("config", ["c1", "c2", "c2tail"])
def test_lu_sign_kernel_matches(config):
    w = synth.make_workload(config)
    s = w.sls
    plan = ha.Plan(w.X.shape[0], s.nv, s.indices, s.offsets, s.dists, s.radius)
    try:
        plan.load(w.X, w.Y)
        plan.execute()
        new = plan.det_signs(lu_kernel=True)
        old = plan.det_signs(lu_kernel=False)
        print(f"[detsign] {config}: {new.sum()} of {len(new)} negative", flush=True)
        assert np.array_equal(new, old)
        # numpy on a sample that holds the largest searchlights and both signs
        rng = np.random.default_rng(0)
        sizes = np.diff(s.offsets)
        which = np.unique(np.concatenate([np.argsort(sizes)[-8:], rng.choice(s.count, 56, replace=False)]))
        assert np.array_equal(new[which], _numpy_negative(plan, which))
        assert 0 < new.sum() < len(new)
    finally:
        plan.close()


def test_reflection_free_pass_gives_det_plus_one():
    """A pass without reflections (the LU-sign kernel, then the fix) leaves every
    local transform with det = +1 on the 10,242-vertex workload."""
    w = synth.make_workload("c1")
    s = w.sls
    plan = ha.Plan(w.X.shape[0], s.nv, s.indices, s.offsets, s.dists, s.radius)
    try:
        plan.load(w.X, w.Y)
        plan.execute(reflection=False)
        assert plan.stats()["n_fallback"] == 0
        assert plan.det_signs(lu_kernel=True).sum() == 0
        assert plan.det_signs(lu_kernel=False).sum() == 0
    finally:
        plan.close()
