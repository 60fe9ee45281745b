"""GPU parity of ridge and searchlight ridge (csrc/ha_ridge.c:13-95,
csrc/ha_searchlight.c:1378-1429) against the CPU FP64 oracle, through the C-ABI.

Tolerances: ha_ridge returns FP64 betas, within 1e-9 relative to max |betas|
(normal equations vs the oracle's SVD, both FP64; the conditioning of these
cases is below 1e4). The searchlight pass stores each searchlight's betas in
FP32 before the FP64 scatter: within 1e-6 relative to max |data|.
"""

import numpy as np
import pytest
import scipy.sparse as sp

import hyperalignment_amd as ha
from hyperalignment_amd import synth
from oracle import ha_oracle as O
from tests import refcases

pytestmark = pytest.mark.gpu

REL_64 = 1e-9
REL_SL = 1e-6


@pytest.fixture(scope="module", autouse=True)
def hip():
    ha.init()
    yield


def rel_err(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_ref_ridge_small_alpha():
    """csrc/ha_test.c:336-362 on the GPU: Y = X B, alpha 1e-6 recovers B to 1e-6."""
    X, Y, B = refcases.ridge_case()
    b = ha.ridge(X, Y, 1e-6)
    assert b.shape == (10, 5)
    assert np.abs(b - B).max() < 1e-6
    assert rel_err(b, O.ridge(X, Y, 1e-6)) < REL_64


@pytest.mark.parametrize("M,N,P,alpha", [(60, 20, 7, 3.0), (30, 12, 20, 1e3), (15, 40, 40, 10.0),
                                         (300, 100, 100, 1e3), (500, 137, 137, 50.0)])
def test_ridge_matches_oracle(M, N, P, alpha):
    rng = np.random.default_rng(M * 7 + N)
    X = rng.standard_normal((M, N))
    Y = rng.standard_normal((M, P))
    print(f"[ridge] M={M} N={N} P={P} alpha={alpha}", flush=True)
    assert rel_err(ha.ridge(X, Y, alpha), O.ridge(X, Y, alpha)) < REL_64


def test_ridge_vec():
    rng = np.random.default_rng(21)
    X = rng.standard_normal((45, 33))
    y = rng.standard_normal(45)
    assert rel_err(ha.ridge(X, y, 4.0), O.ridge(X, y, 4.0)) < REL_64


@pytest.mark.parametrize("M,N,P", [(10, 40, 40), (10, 40, 7), (60, 20, 45), (25, 25, 3)])
def test_ridge_alpha_zero_matches_the_svd_form(M, N, P):
    """alpha = 0 (ha_ridge.c:13-54: d = 1 / s over the thin SVD): with X of full
    rank the betas are the least-squares solution, the minimum-norm one when M
    < N, which the GPU takes from the dual form X^T (X X^T)^-1 Y. P > N runs as
    ceil(P / N) N-column problems, so the padding columns never enter X."""
    rng = np.random.default_rng(M * 131 + N + P)
    X = rng.standard_normal((M, N))
    Y = rng.standard_normal((M, P))
    got = ha.ridge(X, Y, 0.0)
    ref = O.ridge(X, Y, 0.0)
    print(f"[ridge alpha 0] M={M} N={N} P={P}: rel {rel_err(got, ref):.2e}", flush=True)
    assert rel_err(got, ref) < REL_64


def test_ridge_not_positive_definite_is_an_error():
    """alpha = 0 without full rank: a zero column (M >= N, X^T X singular) or a
    zero row (M < N, X X^T singular) leaves an exactly zero pivot, where the
    reference's d = 1 / s divides by a rounding-level singular value."""
    rng = np.random.default_rng(22)
    X = rng.standard_normal((40, 10))
    X[:, 3] = 0.0
    with pytest.raises(ha.HaError):
        ha.ridge(X, rng.standard_normal((40, 10)), 0.0)
    X = rng.standard_normal((10, 40))
    X[4, :] = 0.0
    with pytest.raises(ha.HaError):
        ha.ridge(X, rng.standard_normal((10, 40)), 0.0)
    with pytest.raises(ha.HaError):
        ha.ridge(X, X, -1.0)


def test_searchlight_ridge_c0_weighted():
    w = synth.make_workload("c0")
    s = w.sls
    mat = ha.initialize_sparse_matrix(s.lists(), nv=s.nv)
    wts = O.searchlight_weights_flat(s.indices, s.offsets, s.dists, s.radius, s.nv)
    wl = [wts[s.offsets[i]:s.offsets[i + 1]] for i in range(s.count)]
    print("[ridge] c0 weighted", flush=True)
    got = ha.searchlight_ridge_csc(w.X, w.Y, s.lists(), mat, weights=wl, alpha=1e3)
    ref = O.searchlight_ridge_csc(w.X, w.Y, s.indices, s.offsets, mat.indptr, mat.indices, 1e3, weights=wts)
    assert rel_err(got.data, ref) < REL_SL


def test_searchlight_ridge_accumulates_and_mirrors_python():
    """searchlight.py:176-212: mat0 is copied and accumulated; without mat0 a dense
    array comes back. Unweighted, alpha small, searchlights of up to about 80
    vertices (three 32-column panels) over 40 samples (rank-deficient X_sl)."""
    w = synth.make_workload("c0", n=40, radius=32.0)
    s = w.sls
    mat0 = ha.initialize_sparse_matrix(s.lists(), nv=s.nv)
    mat0.data[:] = np.linspace(-1, 1, mat0.nnz)
    out = ha.searchlight_ridge(w.X, w.Y, s.lists(), mat0=mat0, alpha=0.5)
    assert sp.issparse(out) and np.all(mat0.data == np.linspace(-1, 1, mat0.nnz))
    ref = O.searchlight_ridge_csc(w.X, w.Y, s.indices, s.offsets, mat0.indptr, mat0.indices, 0.5, data0=mat0.data)
    assert rel_err(out.data, ref) < REL_SL
    dense = ha.searchlight_ridge(w.X, w.Y, s.lists(), alpha=0.5)
    ref0 = O.searchlight_ridge_csc(w.X, w.Y, s.indices, s.offsets, mat0.indptr, mat0.indices, 0.5)
    assert isinstance(dense, np.ndarray)
    assert rel_err(dense, O.dense_from_csc(mat0.indptr, mat0.indices, ref0, s.nv)) < REL_SL


def test_searchlight_ridge_other_target_searchlights():
    """sls_Y: the target vertices of each searchlight differ from the source ones."""
    w = synth.make_workload("c0")
    s = w.sls
    perm = np.random.default_rng(5).permutation(s.nv).astype(np.int32)
    idx_y = perm[s.indices]
    lists_y = [idx_y[s.offsets[i]:s.offsets[i + 1]] for i in range(s.count)]
    rows = np.concatenate([np.repeat(s.lists()[i], len(lists_y[i])) for i in range(s.count)])
    cols = np.concatenate([np.tile(lists_y[i], len(s.lists()[i])) for i in range(s.count)])
    mat = sp.csc_matrix((np.zeros(len(rows)), (rows, cols)), shape=(s.nv, s.nv))
    mat.sum_duplicates()
    mat.data[:] = 0.0
    mat.sort_indices()
    got = ha.searchlight_ridge_csc(w.X, w.Y, s.lists(), mat, sls_Y=lists_y, alpha=100.0)
    ref = O.searchlight_ridge_csc(w.X, w.Y, s.indices, s.offsets, mat.indptr, mat.indices, 100.0, idx_y=idx_y)
    assert rel_err(got.data, ref) < REL_SL


def test_plan_ridge_c1():
    """The resident plan at the 10242-vertex, 300-sample size: every nonzero of a
    sample of 24 columns (all the searchlights that touch them) vs the oracle."""
    w = synth.make_workload("c1")
    s = w.sls
    p = ha.Plan(w.X.shape[0], s.nv, s.indices, s.offsets, s.dists, s.radius)
    p.load(w.X, w.Y)
    print("[ridge] c1 plan", flush=True)
    p.ridge(1e3)
    indptr, indices, data = p.csc()
    st = p.stats()
    print(f"[ridge] c1 pass {st['ms_total']:.1f} ms (solve {st['ms_polar']:.1f})", flush=True)
    p.close()
    cols = np.random.default_rng(7).choice(s.nv, 24, replace=False)
    wts = O.searchlight_weights_flat(s.indices, s.offsets, s.dists, s.radius, s.nv)
    ref = O.searchlight_ridge_csc(np.asarray(w.X, np.float64), np.asarray(w.Y, np.float64), s.indices, s.offsets,
                                  indptr, indices, 1e3, weights=wts, which_cols=set(cols.tolist()))
    sel = np.concatenate([np.arange(indptr[c], indptr[c + 1]) for c in cols])
    assert rel_err(data[sel], ref[sel]) < REL_SL


@pytest.mark.parametrize("alpha", [0.0, 5.0])
def test_searchlight_ridge_dual_form(alpha):
    """Searchlights on both sides of nt = 24 (12-70 vertices) in one pass: those
    with k > nt run the dual-form kernel (an nt x nt system), the others the
    k x k one, against the oracle's SVD form. Non-centred data of full row
    rank, so alpha = 0 has the minimum-norm solution everywhere it is dual."""
    rng = np.random.default_rng(77)
    nt, nv = 24, 90
    X = rng.standard_normal((nt, nv)) + 0.5
    Y = rng.standard_normal((nt, nv))
    sizes = [12, 23, 70, 25, 24, 40, 64, 65, 33] if alpha else [25, 33, 40, 64, 65, 70]
    lists = [np.sort(rng.choice(nv, k, replace=False)).astype(np.int32) for k in sizes]
    m = ha.sparse_init(lists, nv)
    ha.searchlight_ridge_csc(X, Y, lists, m, alpha=alpha)
    idx = np.concatenate(lists)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    ref = O.searchlight_ridge_csc(X, Y, idx, off, m.indptr, m.indices, alpha)
    print(f"[ridge dual] alpha {alpha}: rel {rel_err(m.data, ref):.2e}", flush=True)
    assert rel_err(m.data, ref) < REL_SL
