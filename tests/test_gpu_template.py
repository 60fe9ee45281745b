"""Searchlight Procrustes template on the GPU against the FP64 oracle.

The template path keeps FP32 storage of the subjects and of the aligned data,
forms each solve's cross-covariance and polar factor in FP64 (FP64 MFMA) and
the products D Q on FP32 MFMA, so every comparison bounds the max-abs error by
1e-4 times the largest magnitude of the expected values (z-scored or averaged
data, |values| up to about 4). The reference's
own known-answer test (csrc/ha_test.c:503-535) asks 1e-6 of its FP64 path; it
is run here at that relative 1e-4 bar, and so are the hand-derived known
answers of tests/test_oracle_kat.py (one column, exact binary fractions; the
rotation closed forms).
"""

import numpy as np
import pytest

import hyperalignment_amd as ha
from hyperalignment_amd import synth
from oracle import ha_oracle as O
from tests import refcases

pytestmark = pytest.mark.gpu

TOL = 1e-4


def err_ok(got, ref):
    err = np.abs(got - ref).max()
    return err < TOL * max(1.0, np.abs(ref).max()), err


@pytest.fixture(scope="module", autouse=True)
def hip():
    ha.init()
    yield


def test_template_identical_subjects_kat():
    base, dms = refcases.template_case()
    tpl = ha.template_procrustes(dms, True, False, False, 1)
    ok, err = err_ok(tpl, base)
    assert ok, err


@pytest.mark.parametrize("reflection", [True, False])
@pytest.mark.parametrize("zscore,level2,scaling", [(True, 2, False), (False, 1, False), (True, 1, True)])
def test_template_procrustes_matches_oracle(zscore, level2, scaling, reflection):
    rng = np.random.default_rng(5)
    nt, nv, ns = 60, 14, 5
    S = rng.standard_normal((nt, nv))
    dms = []
    for i in range(ns):
        Q, _ = np.linalg.qr(rng.standard_normal((nv, nv)))
        if i % 2:
            Q[:, 0] *= -1.0   # det(Q) alternates, so solves between subjects need the fix
        dms.append(O.zscore_columns(S @ Q + 0.3 * rng.standard_normal((nt, nv))))
    tpl = ha.template_procrustes(dms, reflection, scaling, zscore, level2)
    ref = O.template_procrustes(dms, reflection, scaling, zscore, level2)
    ok, err = err_ok(tpl, ref)
    assert ok, err
    if not reflection:
        # the fix acted: the result differs from the template with reflections
        assert np.abs(tpl - O.template_procrustes(dms, True, scaling, zscore, level2)).max() > 1e-3


@pytest.mark.parametrize("reflection", [True, False])
@pytest.mark.parametrize("nv", [70, 150])
def test_template_procrustes_two_tiles(reflection, nv):
    """70 columns (kp 96: the FP32 products D Q in two 64-wide tiles) and 150
    (kp 160: the FP64 products of the cross-covariance and the polar iteration
    in 2 x 2 tiles of 128, the symmetric ones forming the upper tiles and
    mirroring them), four subjects, with and without reflections, against the
    oracle."""
    rng = np.random.default_rng(9)
    nt, ns = 200, 4
    S = rng.standard_normal((nt, nv))
    dms = []
    for i in range(ns):
        Q, _ = np.linalg.qr(rng.standard_normal((nv, nv)))
        if i % 2:
            Q[:, 0] *= -1.0
        dms.append(O.zscore_columns(S @ Q + 0.3 * rng.standard_normal((nt, nv))))
    tpl = ha.template_procrustes(dms, reflection, False, True, 2)
    ref = O.template_procrustes(dms, reflection, False, True, 2)
    ok, err = err_ok(tpl, ref)
    print(f"[template nv {nv}] reflection {reflection}: max |diff| {err:.2e}", flush=True)
    assert ok, err
    if not reflection:
        assert np.abs(tpl - O.template_procrustes(dms, True, False, True, 2)).max() > 1e-3


@pytest.mark.parametrize("weighted,reflection", [(True, True), (False, True), (True, False)])
def test_searchlight_template_small_mesh(weighted, reflection):
    """Small mesh (order 3, 30 mm: 642 searchlights of 13-18 vertices), with
    and without reflections (ha_template.c:40,80 pass isReflection to every
    Procrustes solve; without them each gets det = +1)."""
    dms, sls = synth.template_subjects(n_subj=4, n=80, order=3, radius=30.0, seed=3)
    d = sls.dists if weighted else None
    tpl = ha.searchlight_template_flat(dms, sls.indices, sls.offsets, d, sls.radius, reflection=reflection,
                                       level2_iter=2)
    ref = O.searchlight_template(dms, sls.indices, sls.offsets, d, sls.radius, reflection=reflection,
                                 level2_iter=2, n_jobs=8)
    ok, err = err_ok(tpl, ref)
    assert ok, err


@pytest.mark.parametrize("ns,expected", [(3, "TPL_3"), (4, "TPL_4")])
def test_template_hand_kat(ns, expected):
    from tests import test_oracle_kat as K
    dms = [np.array(x, dtype=float).reshape(3, 1) for x in K.TPL_X[:ns]]
    tpl = ha.template_procrustes(dms, True, True, False, 1)
    ok, err = err_ok(tpl.ravel(), np.array(getattr(K, expected)))
    assert ok, err


@pytest.mark.parametrize("zscore", [False, True])
def test_template_rotations_closed_form(zscore):
    from tests import test_oracle_kat as K
    dms, S, Rs, a, c = K.rotation_subjects()
    tpl = ha.template_procrustes(dms, True, True, zscore, 2)
    ref = S @ Rs[0] / np.sqrt(c / S.shape[0]) if zscore else a[0] * S @ Rs[0]
    ok, err = err_ok(tpl, ref)
    assert ok, err


# The 16-subject workload's local cross-covariances are ill-conditioned (sigma_min / sigma_max
# about 2.5e-4 at 20 mm, n = 400, and down to 1.8e-8 in a few solves), and each
# searchlight chains 47 Procrustes solves. With FP32 cross-covariances and
# polar iterations one searchlight's local template was 0.89 off, and these 16
# columns 2.0e-4 of the largest magnitude; with both in FP64 they measure
# 3.5e-7 absolute (1.1e-7 relative). Bar: max-abs 1e-4 times the largest
# magnitude of the expected values, as for the other template tests.
TOL_C4 = 1e-4


def test_template_c4_sampled_columns():
    """The 16-subject workload (order-5 mesh, n=400, 20 mm, level2 = 2): 16
    template columns (vertices of one searchlight) against the oracle over
    every searchlight holding any of them."""
    import sys
    import time
    t0 = time.time()
    dms, sls = synth.template_subjects(n_subj=16, n=400, order=5, radius=20.0, seed=0)
    print(f"[c4] data {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    tpl = ha.searchlight_template_flat(dms, sls.indices, sls.offsets, sls.dists, sls.radius, level2_iter=2)
    print(f"[c4] GPU template {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    assert np.all(np.isfinite(tpl))
    lists = sls.lists()
    cols = [int(c) for c in lists[7777][::6][:16]]
    assert len(cols) == 16
    colset = set(cols)
    which = [s for s in range(sls.count) if colset & set(lists[s].tolist())]
    ref = O.searchlight_template(dms, sls.indices, sls.offsets, sls.dists, sls.radius, level2_iter=2,
                                 which=which, n_jobs=16)
    err = np.abs(tpl[:, cols] - ref[:, cols]).max()
    mag = np.abs(ref[:, cols]).max()
    print(f"[c4] oracle over {len(which)} searchlights {time.time() - t0:.1f} s: max-abs {err:.3e}, "
          f"relative to max |value| ({mag:.2f}) {err / mag:.3e}", file=sys.stderr, flush=True)
    assert err < TOL_C4 * max(1.0, mag), err


@pytest.mark.parametrize("nv", [1, 2, 3, 5])
def test_template_few_columns_without_reflections(nv):
    """Fewer columns than the reflection fix's four iteration vectors: its
    Rayleigh-Ritz step then uses only the first k of them (the others hold no
    direction of their own), and every solve still ends with det = +1."""
    rng = np.random.default_rng(40 + nv)
    nt, ns = 30, 4
    S = rng.standard_normal((nt, nv))
    dms = []
    for i in range(ns):
        Q, _ = np.linalg.qr(rng.standard_normal((nv, nv)))
        if i % 2:
            Q[:, 0] *= -1.0
        dms.append(O.zscore_columns(S @ Q + 0.3 * rng.standard_normal((nt, nv))))
    tpl = ha.template_procrustes(dms, False, False, True, 2)
    ref = O.template_procrustes(dms, False, False, True, 2)
    ok, err = err_ok(tpl, ref)
    print(f"[template nv {nv}] without reflections: max |diff| {err:.2e}", flush=True)
    assert ok, err


@pytest.mark.parametrize("reflection", [True, False])
@pytest.mark.parametrize("zscore,level2", [(True, 2), (False, 1)])
def test_template_fewer_samples_than_columns(reflection, zscore, level2):
    """nt = 10 samples, 14 columns: every cross-covariance of the chain is
    singular (rank <= nt - 1), below what the polar iteration's schedules
    reach, so each solve takes the FP64 Jacobi SVD, which completes U over the
    zero singular values. With z-scored subjects the null space of A is that
    of D_s, so D_s Q (and the template) does not depend on the completion and
    matches the oracle's LAPACK solve; without reflections the flipped pair is
    a zero one and changes nothing either."""
    rng = np.random.default_rng(21)
    nt, nv, ns = 10, 14, 4
    S = rng.standard_normal((nt, nv))
    dms = [O.zscore_columns(S @ np.linalg.qr(rng.standard_normal((nv, nv)))[0] + 0.3 * rng.standard_normal((nt, nv)))
           for _ in range(ns)]
    tpl = ha.template_procrustes(dms, reflection, False, zscore, level2)
    ref = O.template_procrustes(dms, reflection, False, zscore, level2)
    ok, err = err_ok(tpl, ref)
    print(f"[template nt < nv] reflection {reflection} zscore {zscore}: max |diff| {err:.2e}", flush=True)
    assert ok, err


@pytest.mark.parametrize("reflection", [True, False])
def test_searchlight_template_few_samples(reflection):
    """Small mesh (order 3, 30 mm: searchlights of 13-18 vertices) with 15
    samples, so some searchlights have k >= nt (singular chains, the Jacobi
    solve) and some k < nt (the polar iteration), against the oracle."""
    dms, sls = synth.template_subjects(n_subj=4, n=15, order=3, radius=30.0, seed=4)
    sizes = np.diff(sls.offsets)
    assert (sizes >= 15).any() and (sizes < 15).any()
    tpl = ha.searchlight_template_flat(dms, sls.indices, sls.offsets, sls.dists, sls.radius, reflection=reflection,
                                       level2_iter=2)
    ref = O.searchlight_template(dms, sls.indices, sls.offsets, sls.dists, sls.radius, reflection=reflection,
                                 level2_iter=2, n_jobs=8)
    ok, err = err_ok(tpl, ref)
    print(f"[template mesh nt 15] reflection {reflection}: max |diff| {err:.2e}", flush=True)
    assert ok, err


@pytest.mark.parametrize("reflection", [True, False])
def test_template_subject_with_zero_column(reflection):
    """A subject with an all-zero column (a vertex without signal): its
    cross-covariances have an exactly zero singular value, which the FP64
    iteration cannot finish, so its solves take the Jacobi solve. The zero
    column is in the null space of D_s, so D_s Q does not depend on the
    completion, and the template matches the oracle (the reference's SVD
    gives a valid T there too, where this path used to report kHaErrorSvd)."""
    rng = np.random.default_rng(33)
    nt, nv, ns = 50, 12, 4
    S = rng.standard_normal((nt, nv))
    dms = [O.zscore_columns(S @ np.linalg.qr(rng.standard_normal((nv, nv)))[0] + 0.3 * rng.standard_normal((nt, nv)))
           for _ in range(ns)]
    dms[2][:, 5] = 0.0
    tpl = ha.template_procrustes(dms, reflection, False, True, 2)
    ref = O.template_procrustes(dms, reflection, False, True, 2)
    ok, err = err_ok(tpl, ref)
    print(f"[template zero column] reflection {reflection}: max |diff| {err:.2e}", flush=True)
    assert ok, err


@pytest.mark.parametrize("reflection", [True, False])
def test_template_fewer_samples_than_columns_with_scaling(reflection):
    """isScaling on singular chains (nt = 10, 14 columns): the Jacobi solve's
    sum of singular values (less twice a flipped one, zero here) scales each
    D_s Q as the reference's sum(s) / sum_j var, against the oracle."""
    rng = np.random.default_rng(23)
    nt, nv, ns = 10, 14, 4
    S = rng.standard_normal((nt, nv))
    dms = [O.zscore_columns(S @ np.linalg.qr(rng.standard_normal((nv, nv)))[0] + 0.3 * rng.standard_normal((nt, nv)))
           for _ in range(ns)]
    tpl = ha.template_procrustes(dms, reflection, True, True, 1)
    ref = O.template_procrustes(dms, reflection, True, True, 1)
    ok, err = err_ok(tpl, ref)
    print(f"[template nt < nv, scaling] reflection {reflection}: max |diff| {err:.2e}", flush=True)
    assert ok, err


def test_last_template_stats():
    """ha_hip_last_template_stats reports the last flat call: its kernel time,
    the polar steps of its solves (at least one per solve) and their FP64 flops
    (at least the cross-covariances, 2 n k^2 per solve)."""
    dms, sls = synth.template_subjects(n_subj=4, n=80, order=3, radius=30.0, seed=3)
    ha.searchlight_template_flat(dms, sls.indices, sls.offsets, sls.dists, sls.radius, level2_iter=2)
    st = ha.last_template_stats()
    k = np.diff(sls.offsets).astype(np.float64)
    n_solves = 3 + 2 * 4
    assert st.ms_polar > 0
    assert st.steps_total >= n_solves * sls.count
    assert st.gemm_flops > (n_solves * 2.0 * 80 * k * k).sum()
    assert st.n_unconverged == 0
