"""Known answers derived by hand from the reference's C code, pinning the parts
of the oracle that its other checks leave open: the direction of the searchlight
weights in the drivers and the template recurrence.

Driver (csrc/ha_searchlight.c:1072-1085 dense, csrc/ha_sparse.c:128-142 via
:193-198 CSC): T_out[sl[i], sl[j]] += T[i, j] * w[i], the weight of the ROW
vertex. Weighting columns instead (src/hyperalignment/searchlight.py:125)
changes entries (1, 0), (1, 2), (2, 1) and (0, 1) of the case below.

Case: three vertices, X = I (three samples), Y = B with
  B = [[3, -4, 0], [4, 3, 4], [0, -4, 3]],
searchlights s0 = {0, 1} (distances 0, 4) and s1 = {1, 2} (distances 3, 0),
radius 10. With X = I, A_s = B[s, s]: s0 -> [[3, -4], [4, 3]] = 5 R(+),
s1 -> [[3, 4], [-4, 3]] = 5 R(-), so T_0 = [[.6, -.8], [.8, .6]] and
T_1 = [[.6, .8], [-.8, .6]] (neither symmetric). Raw weights (r - d)/r:
s0 (1, .6), s1 (.7, 1); per-vertex sums 1, 1.3, 1; normalised s0 (1, 6/13),
s1 (7/13, 1) (csrc/ha_searchlight.c:970-1001). Row-weighted sum:
  T_out = [[.6,       -.8,           0      ],
           [.8*6/13,   .6*(6+7)/13,  .8*7/13],
           [0,        -.8,           .6     ]].

Template (csrc/ha_template.c:14-114) with one column (v = 1), isScaling, no
z-scores: for mean-zero x, ha_procrustes gives T = sign(x.ref) |x.ref| /
(n var(x)) = x.ref / |x|^2, so each alignment is the projection of the
reference onto x. Subjects x0 = (1, -1, 0), x1 = (1, 0, -1), x2 = (0, 1, -1),
x3 = (2, -1, -1). Three subjects, level2_iter = 1:
  sequential: common = x0; a1 = x1 (x1.x0)/2 = (1/2, 0, -1/2),
    common = (x0 + a1)/2 = (3/4, -1/2, -1/4); a2 = x2 (-1/4)/2 = (0, -1/8, 1/8)
  level 2, sum = x0 + a1 + a2 = (3/2, -9/8, -3/8) taken once before the loop,
    ref_s = (sum - a_s)/(ns - 1):
    ref_0 = (1/4, -1/16, -3/16)  -> a0 = x0 (5/16)/2 = (5/32, -5/32, 0)
    ref_1 = (1/2, -9/16, 1/16)   -> a1 = x1 (7/16)/2 = (7/32, 0, -7/32)
    ref_2 = (3/4, -1/2, -1/4)    -> a2 = x2 (-1/4)/2 = (0, -1/8, 1/8)
  template = (a0 + a1 + a2)/3 = (1/8, -3/32, -1/32).
Dividing by ns instead of ns - 1 scales the template by 2/3; updating the sum
inside the loop changes ref_1 and ref_2. Four subjects (x0..x3), same steps:
  sequential: a1 = (1/2, 0, -1/2), common (3/4, -1/2, -1/4); a2 = (0, -1/8, 1/8),
    common = (3/8, -5/16, -1/16); a3 = x3 (9/8)/6 = (3/8, -3/16, -3/16)
    (a running mean 1/(s+1) instead of the halving changes a3);
  sum = (15/8, -21/16, -9/16); a0 = (19/96, -19/96, 0), a1 = (23/96, 0, -23/96),
    a2 = (0, -1/12, 1/12), a3 = (1/2, -1/4, -1/4);
  template = (15/64, -17/128, -13/128).

Rotations: subjects D_s = a_s S R_s with S^T S = c I (mean-zero columns) and
orthogonal R_s. Procrustes of D_s onto b S R' is (b / a_s) R_s^T R' with
scaling, so every alignment reproduces its reference exactly: the running
common space stays a_0 S R_0 and, with level2_iter = 2 and no z-scores, the
template is a_0 S R_0; with z-scores every aligned subject becomes S R_0 /
sqrt(c / n) and so does the template.
"""

import numpy as np

from oracle import ha_oracle as O

KAT_X = np.eye(3)
KAT_Y = np.array([[3.0, -4.0, 0.0], [4.0, 3.0, 4.0], [0.0, -4.0, 3.0]])
KAT_IDX = np.array([0, 1, 1, 2], dtype=np.int32)
KAT_OFF = np.array([0, 2, 4], dtype=np.int32)
KAT_DIST = np.array([0.0, 4.0, 3.0, 0.0])
KAT_RADIUS = 10.0
KAT_T = np.array([[0.6, -0.8, 0.0],
                  [0.8 * 6 / 13, 0.6, 0.8 * 7 / 13],
                  [0.0, -0.8, 0.6]])

TPL_X = [(1, -1, 0), (1, 0, -1), (0, 1, -1), (2, -1, -1)]
TPL_3 = (1 / 8, -3 / 32, -1 / 32)
TPL_4 = (15 / 64, -17 / 128, -13 / 128)


def rotation_subjects(ns=4, n=12, v=5, seed=0):
    rng = np.random.default_rng(seed)
    Z = rng.standard_normal((n, v))
    Z -= Z.mean(axis=0)
    S, _ = np.linalg.qr(Z)          # orthonormal, mean-zero columns
    c = 2.5
    S *= np.sqrt(c)
    Rs = [np.linalg.qr(rng.standard_normal((v, v)))[0] for _ in range(ns)]
    a = [1.0, 0.5, 2.0, 1.5][:ns]
    return [a[i] * S @ Rs[i] for i in range(ns)], S, Rs, a, c


def test_driver_row_weights_dense():
    T = O.searchlight_procrustes_dense(KAT_X, KAT_Y, KAT_IDX, KAT_OFF, KAT_DIST, KAT_RADIUS)
    assert np.abs(T - KAT_T).max() < 1e-14


def test_driver_row_weights_csc():
    indptr, indices = O.sparse_init(KAT_IDX, KAT_OFF, 3)
    w = O.searchlight_weights_flat(KAT_IDX, KAT_OFF, KAT_DIST, KAT_RADIUS, 3)
    data = O.searchlight_procrustes_csc(KAT_X, KAT_Y, KAT_IDX, KAT_OFF, indptr, indices, weights=w)
    assert np.abs(O.dense_from_csc(indptr, indices, data, 3) - KAT_T).max() < 1e-14
    # the two zero corners are outside the pattern
    assert indptr.tolist() == [0, 2, 5, 7]


def test_template_three_subjects_by_hand():
    dms = [np.array(x, dtype=float).reshape(3, 1) for x in TPL_X[:3]]
    tpl = O.template_procrustes(dms, reflection=True, scaling=True, zscore_common=False, level2_iter=1)
    assert np.abs(tpl.ravel() - np.array(TPL_3)).max() < 1e-15


def test_template_four_subjects_by_hand():
    dms = [np.array(x, dtype=float).reshape(3, 1) for x in TPL_X]
    tpl = O.template_procrustes(dms, reflection=True, scaling=True, zscore_common=False, level2_iter=1)
    assert np.abs(tpl.ravel() - np.array(TPL_4)).max() < 1e-15


def test_template_rotations_closed_form():
    dms, S, Rs, a, c = rotation_subjects()
    tpl = O.template_procrustes(dms, reflection=True, scaling=True, zscore_common=False, level2_iter=2)
    assert np.abs(tpl - a[0] * S @ Rs[0]).max() < 1e-12
    n = S.shape[0]
    tpl = O.template_procrustes(dms, reflection=True, scaling=True, zscore_common=True, level2_iter=2)
    assert np.abs(tpl - S @ Rs[0] / np.sqrt(c / n)).max() < 1e-12


def _newton(A, iters=60):
    """X <- (X + X^-T)/2 from X = A, csrc/ha_procrustes.c:280-306, in FP64."""
    X = np.array(A, dtype=np.float64)
    for _ in range(iters):
        Xn = 0.5 * (X + np.linalg.inv(X).T)
        if np.linalg.norm(Xn - X) <= 1e-15 * np.linalg.norm(Xn):
            return Xn
        X = Xn
    return X


def test_polar_newton_is_the_newton_limit():
    rng = np.random.default_rng(8)
    for n in (3, 7, 20):
        A = rng.standard_normal((n, n))
        assert np.abs(O.polar_newton(A) - _newton(A)).max() < 1e-12


def test_polar_newton_reflection_negates_last_column():
    """det(A) < 0 and no reflections: the Newton factor with its last column
    negated (csrc/ha_procrustes.c:328-331), not the SVD rank-1 flip that
    ha_procrustes applies (csrc/ha_procrustes.c:56-95)."""
    rng = np.random.default_rng(9)
    A = rng.standard_normal((6, 6))
    if np.linalg.det(A) > 0:
        A[:, 0] *= -1.0
    T = O.polar_newton(A, reflection=False)
    ref = _newton(A)
    ref[:, -1] *= -1.0
    assert np.abs(T - ref).max() < 1e-12
    assert np.linalg.det(T) > 0
    # positive determinant: untouched
    B = A.copy()
    B[:, 0] *= -1.0
    assert np.abs(O.polar_newton(B, reflection=False) - _newton(B)).max() < 1e-12


def test_polar_newton_scaling_closed_form():
    """Y = c X R (R orthogonal, X mean-zero columns): A = X^T Y = c X^T X R, the
    polar factor is R, trace(R^T A) = c trace(X^T X) = c M sum var(X), so the
    scaled factor is c R."""
    rng = np.random.default_rng(10)
    X = rng.standard_normal((40, 5))
    X -= X.mean(axis=0)
    R = np.linalg.qr(rng.standard_normal((5, 5)))[0]
    A = X.T @ (2.5 * X @ R)
    assert np.abs(O.polar_newton(A, scaling=True, X=X) - 2.5 * R).max() < 1e-12
