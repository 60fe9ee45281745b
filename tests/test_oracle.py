"""Pins the CPU FP64 oracle before it is used to judge the GPU path.

Checks come from the reference's own C tests (csrc/ha_test.c), closed forms,
invariants and an independent algorithm (eigendecomposition instead of SVD),
chosen so that a dropped term, a wrong index or a transposed operand fails.
"""

import numpy as np
import pytest
import scipy.linalg as sla

from hyperalignment_amd import synth
from oracle import ha_oracle as O
from tests import refcases


def polar_eigh(A):
    """Polar factor by A (A^T A)^{-1/2} via eigh: an algorithm independent of SVD."""
    w, V = np.linalg.eigh(A.T @ A)
    return A @ (V * (1.0 / np.sqrt(w))) @ V.T


# ---- reference known-answer tests ----

def test_procrustes_identity():
    X = refcases.identity_case()
    T = O.procrustes(X, X, True, False)
    assert np.abs(np.abs(T) - np.eye(20)).max() < 1e-8


def test_procrustes_known_rotation():
    X, Y, R = refcases.rotation_case()
    T = O.procrustes(X, Y, True, False)
    assert np.abs(X @ T - Y).max() < 1e-8
    assert np.abs(T - R).max() < 1e-8


def test_procrustes_no_reflection():
    X, Y = refcases.no_reflection_case()
    T = O.procrustes(X, Y, False, False)
    assert np.linalg.det(T) > 0
    # the unconstrained solution of this case is a reflection, so the fix is exercised
    assert np.linalg.det(O.procrustes(X, Y, True, False)) < 0
    assert np.abs(T.T @ T - np.eye(10)).max() < 1e-10


def test_sparse_init_toy():
    indptr, indices = O.sparse_init_pairs(refcases.SPARSE_TOY, 5)
    assert indptr[-1] == 21
    sl = synth.toy_searchlights(refcases.SPARSE_TOY, nv=5)
    p2, i2 = O.sparse_init(sl.indices, sl.offsets, 5)
    assert np.array_equal(indptr, p2) and np.array_equal(indices, i2)


def test_scatter_identity():
    # csrc/ha_test.c:404-435: identity scattered into a 3 x 3 pattern
    sl = synth.toy_searchlights([[0, 1, 2]], nv=3)
    indptr, indices = O.sparse_init(sl.indices, sl.offsets, 3)
    X = np.random.default_rng(0).standard_normal((20, 3))
    data = O.searchlight_procrustes_csc(X, X, sl.indices, sl.offsets, indptr, indices)
    D = O.dense_from_csc(indptr, indices, data, 3)
    assert np.abs(D - np.eye(3)).max() < 1e-12


def test_weights_uniform_toy():
    sl = synth.toy_searchlights(refcases.WEIGHTS_TOY, nv=6)
    w = O.searchlight_weights_flat(sl.indices, sl.offsets, None, 0.0, 6)
    sums = np.zeros(6)
    np.add.at(sums, sl.indices, w)
    assert np.abs(sums - 1.0).max() < 1e-12


def test_zscore():
    X = O.zscore_columns(refcases.zscore_case())
    assert np.abs(X.mean(axis=0)).max() < 1e-12
    assert np.abs((X * X).mean(axis=0) - X.mean(axis=0) ** 2 - 1.0).max() < 1e-10


def test_template_identical_subjects():
    base, dms = refcases.template_case()
    tpl = O.template_procrustes(dms, True, False, False, 1)
    assert np.abs(tpl - base).max() < 1e-6


# ---- weights: bit-exact ordering ----

@pytest.mark.parametrize("uniform", [True, False])
def test_weights_bitexact_vs_loop(uniform):
    w = synth.make_workload("c0")
    s = w.sls
    d = None if uniform else s.dists
    a = O.searchlight_weights_flat(s.indices, s.offsets, d, s.radius, s.nv)
    b = O.searchlight_weights_loop(s.indices, s.offsets, d, s.radius, s.nv)
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    sums = np.zeros(s.nv)
    np.add.at(sums, s.indices, a)
    assert np.abs(sums - 1).max() < 1e-12


# ---- pattern: scalable union == reference pair enumeration ----

def test_pattern_union_equals_pairs_c0():
    w = synth.make_workload("c0")
    s = w.sls
    p1, i1 = O.sparse_init_pairs(s.lists(), s.nv)
    p2, i2 = O.sparse_init(s.indices, s.offsets, s.nv)
    assert np.array_equal(p1.astype(np.int64), p2) and np.array_equal(i1, i2)
    # symmetric, diagonal present, rows sorted
    assert all(np.all(np.diff(i2[p2[c]:p2[c + 1]]) > 0) for c in range(0, s.nv, 97))


def test_pattern_brute_force_random():
    rng = np.random.default_rng(5)
    nv = 40
    lists = [np.sort(rng.choice(nv, rng.integers(1, 9), replace=False)) for _ in range(25)]
    M = np.zeros((nv, nv), dtype=bool)
    for l in lists:
        M[np.ix_(l, l)] = True
    p, i = O.sparse_init_pairs(lists, nv)
    for c in range(nv):
        assert np.array_equal(i[p[c]:p[c + 1]], np.nonzero(M[:, c])[0])


# ---- local Procrustes: independent algorithm, argmin, scaling ----

def test_procrustes_matches_eigh_polar():
    rng = np.random.default_rng(1)
    X = rng.standard_normal((60, 25))
    Y = X @ np.linalg.qr(rng.standard_normal((25, 25)))[0] + 0.5 * rng.standard_normal((60, 25))
    T = O.procrustes(X, Y)
    assert np.abs(T - polar_eigh(X.T @ Y)).max() < 1e-10


def test_procrustes_is_argmin():
    rng = np.random.default_rng(2)
    X = rng.standard_normal((40, 12))
    Y = rng.standard_normal((40, 12))
    T = O.procrustes(X, Y)
    base = np.linalg.norm(X @ T - Y)
    for _ in range(20):
        S = rng.standard_normal((12, 12))
        R = sla.expm(1e-3 * (S - S.T))
        assert np.linalg.norm(X @ T @ R - Y) >= base - 1e-12


def test_procrustes_scaling_closed_form():
    rng = np.random.default_rng(3)
    X = rng.standard_normal((80, 9))
    X -= X.mean(axis=0)
    R = np.linalg.qr(rng.standard_normal((9, 9)))[0]
    T = O.procrustes(X, 2.5 * X @ R, True, True)
    assert np.abs(T - 2.5 * R).max() < 1e-10


def test_det_sign_matches_det():
    rng = np.random.default_rng(4)
    for _ in range(10):
        A = rng.standard_normal((7, 7))
        assert O.det_sign(A) == np.sign(np.linalg.det(A))


# ---- searchlight drivers: closed forms ----

def test_driver_identity_and_negation():
    w = synth.make_workload("c0")
    s = w.sls
    T = O.searchlight_procrustes_dense(w.X, w.X, s.indices, s.offsets, s.dists, s.radius)
    assert np.abs(T - np.eye(s.nv)).max() < 1e-10
    T = O.searchlight_procrustes_dense(w.X, -w.X, s.indices, s.offsets, s.dists, s.radius)
    assert np.abs(T + np.eye(s.nv)).max() < 1e-10


def test_driver_single_global_searchlight():
    rng = np.random.default_rng(6)
    X = rng.standard_normal((30, 8))
    Y = rng.standard_normal((30, 8))
    sl = synth.toy_searchlights([list(range(8))], nv=8)
    T = O.searchlight_procrustes_dense(X, Y, sl.indices, sl.offsets, None, 0.0)
    assert np.abs(T - polar_eigh(X.T @ Y)).max() < 1e-10


def test_driver_csc_equals_dense_c0():
    w = synth.make_workload("c0")
    s = w.sls
    indptr, indices = O.sparse_init(s.indices, s.offsets, s.nv)
    wts = O.searchlight_weights_flat(s.indices, s.offsets, s.dists, s.radius, s.nv)
    data = O.searchlight_procrustes_csc(w.X, w.Y, s.indices, s.offsets, indptr, indices, weights=wts)
    D = O.searchlight_procrustes_dense(w.X, w.Y, s.indices, s.offsets, s.dists, s.radius)
    assert np.array_equal(O.dense_from_csc(indptr, indices, data, s.nv), D)


def test_searchlight_template_single_searchlight_is_template():
    """One searchlight over every vertex (weight 1 each) gives ha_template.c's template."""
    rng = np.random.default_rng(4)
    dms = [O.zscore_columns(rng.standard_normal((40, 9))) for _ in range(4)]
    idx = np.arange(9, dtype=np.int32)
    off = np.array([0, 9], dtype=np.int32)
    a = O.searchlight_template(dms, idx, off, None, 0.0, level2_iter=2)
    b = O.template_procrustes(dms, True, False, True, 2)
    assert np.abs(a - b).max() < 1e-12


def test_apply_transform_matches_scipy():
    import scipy.sparse as sp
    rng = np.random.default_rng(8)
    W = sp.random(30, 25, density=0.2, random_state=3, format="csc")
    X = rng.standard_normal((17, 30))
    got = O.apply_transform(X, W.indptr, W.indices, W.data, 25)
    assert np.abs(got - X @ W).max() < 1e-13


def test_vertex_correlation_is_pearson():
    rng = np.random.default_rng(9)
    A = rng.standard_normal((50, 6))
    B = A + rng.standard_normal((50, 6))
    r = O.vertex_correlation(A, B)
    ref = [np.corrcoef(A[:, j], B[:, j])[0, 1] for j in range(6)]
    assert np.abs(r - ref).max() < 1e-13
    assert np.isnan(O.vertex_correlation(np.ones((5, 1)), B[:5, :1]))[0]


# ---- ridge ----

def test_ridge_recovers_true_betas():
    """csrc/ha_test.c:336-362: Y = X B, alpha 1e-6 -> max |betas - B| < 1e-6."""
    X, Y, B = refcases.ridge_case()
    assert np.abs(O.ridge(X, Y, 1e-6) - B).max() < 1e-6


@pytest.mark.parametrize("M,N,P,alpha", [(60, 20, 7, 3.0), (15, 40, 40, 10.0), (30, 12, 20, 1e3)])
def test_ridge_normal_equations(M, N, P, alpha):
    """(X^T X + alpha I) betas = X^T Y, including fewer samples than features."""
    rng = np.random.default_rng(M + N + P)
    X = rng.standard_normal((M, N))
    Y = rng.standard_normal((M, P))
    b = O.ridge(X, Y, alpha)
    lhs = (X.T @ X + alpha * np.eye(N)) @ b
    assert np.abs(lhs - X.T @ Y).max() < 1e-10 * max(1.0, np.abs(X.T @ Y).max())


def test_ridge_limits():
    rng = np.random.default_rng(12)
    X = rng.standard_normal((40, 9))
    Y = rng.standard_normal((40, 4))
    # alpha = 0 and full column rank: least squares
    assert np.abs(O.ridge(X, Y, 0.0) - np.linalg.lstsq(X, Y, rcond=None)[0]).max() < 1e-12
    # alpha -> inf: alpha betas -> X^T Y
    a = 1e12
    assert np.abs(a * O.ridge(X, Y, a) - X.T @ Y).max() < 1e-6 * np.abs(X.T @ Y).max()
    # one response vector (csrc/ha_ridge.c:56-95) is a column of the matrix form
    assert np.abs(O.ridge(X, Y[:, 2], 5.0) - O.ridge(X, Y, 5.0)[:, 2]).max() < 1e-14


def test_searchlight_ridge_brute_force():
    """mat[np.ix_(sl, sl)] += betas_s * w_s[:, None] over the searchlights in order
    (csrc/ha_searchlight.c:1378-1429), through the CSC pattern."""
    rng = np.random.default_rng(13)
    lists = [[0, 1, 2, 5], [1, 2, 3], [3, 4, 5, 6, 0], [6, 2]]
    nv = 7
    sl = synth.toy_searchlights(lists, nv=nv)
    X = rng.standard_normal((12, nv))
    Y = rng.standard_normal((12, nv))
    wts = rng.uniform(0.2, 1.0, size=int(sl.offsets[-1]))
    indptr, indices = O.sparse_init(sl.indices, sl.offsets, nv)
    data = O.searchlight_ridge_csc(X, Y, sl.indices, sl.offsets, indptr, indices, 2.5, weights=wts)
    D = np.zeros((nv, nv))
    for s, l in enumerate(lists):
        b = np.linalg.solve(X[:, l].T @ X[:, l] + 2.5 * np.eye(len(l)), X[:, l].T @ Y[:, l])
        D[np.ix_(l, l)] += b * wts[sl.offsets[s]:sl.offsets[s + 1], None]
    assert np.abs(O.dense_from_csc(indptr, indices, data, nv) - D).max() < 1e-12
