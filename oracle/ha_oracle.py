"""CPU FP64 oracle for searchlight Procrustes -- TEST INFRASTRUCTURE ONLY.

A restatement, in numpy/scipy, of the reference's CPU FP64 path. It is the
checker for the HIP backend: only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import it, and never as the thing
that is measured or shipped.

Every function cites the reference lines it follows (paths relative to the
reference repository). SVDs use LAPACK ``dgesdd`` with a ``dgesvd`` fallback
through scipy, the same routines the reference calls
(csrc/ha_searchlight.c:593-605, csrc/ha_linalg.c:53-89, src/hyperalignment/linalg.py:45-51).

How each function is pinned (tests/test_oracle.py):
  searchlight_weights_flat  bit-exact against a pure-Python transcription of the
                            reference loop, plus the sum-to-one invariant
                            (csrc/ha_test.c:439-470)
  sparse_init_pairs         reference toy case nnz == 21 (csrc/ha_test.c:369-402)
                            and brute-force dense marking
  sparse_init               equals sparse_init_pairs on every test mesh
  procrustes                identity, Givens 0.7 rad rotation and det > 0 cases of
                            csrc/ha_test.c:234-332; argmin property against
                            random orthogonal perturbations; scaling closed form
  searchlight_procrustes_*  brute-force dense accumulation on tiny inputs and the
                            CSC/dense agreement
  polar_newton              the reference's own Newton iteration run to convergence
                            (csrc/ha_procrustes.c:271-306); reflection and scaling
                            rules checked on constructed cases
  template_procrustes       identical subjects give the input back
                            (csrc/ha_test.c:503-535)
  ridge                     recovers exact betas at alpha 1e-6 (csrc/ha_test.c:336-362);
                            normal equations (X^T X + alpha I) B = X^T Y, also with
                            fewer samples than features; alpha -> inf limit X^T Y / alpha;
                            alpha = 0 equals least squares
  searchlight_ridge_csc     brute-force dense accumulation on a tiny mesh
"""

from __future__ import annotations

import numpy as np
import scipy.linalg as sla


# ---------------------------------------------------------------------------
# weights
# ---------------------------------------------------------------------------

def searchlight_weights_flat(idx, offsets, dists, radius, nv):
    """Flat weights as csrc/ha_searchlight.c:970-1001 (also :36-92).

    Per-vertex sums accumulate in flat order k = 0..total-1 (np.add.at is
    unbuffered and applies in index order), so every weight is bit-identical
    to the reference's sequential loop.
    """
    idx = np.asarray(idx, dtype=np.int64)
    wsum = np.zeros(int(nv), dtype=np.float64)
    if dists is None:
        np.add.at(wsum, idx, 1.0)
        return 1.0 / wsum[idx]
    w = (radius - np.asarray(dists, dtype=np.float64)) / radius
    np.add.at(wsum, idx, w)
    return w / wsum[idx]


def searchlight_weights_loop(idx, offsets, dists, radius, nv):
    """Pure-Python transcription of csrc/ha_searchlight.c:970-1001 (small inputs)."""
    total = int(offsets[-1])
    wsum = [0.0] * int(nv)
    out = [0.0] * total
    if dists is None:
        for k in range(total):
            wsum[int(idx[k])] += 1.0
        for k in range(total):
            out[k] = 1.0 / wsum[int(idx[k])]
    else:
        for k in range(total):
            wsum[int(idx[k])] += (radius - float(dists[k])) / radius
        for k in range(total):
            w = (radius - float(dists[k])) / radius
            out[k] = w / wsum[int(idx[k])]
    return np.array(out)


# ---------------------------------------------------------------------------
# CSC pattern
# ---------------------------------------------------------------------------

def sparse_init_pairs(sl_lists, nv):
    """csrc/ha_sparse.c:31-112: all (sl[i], sl[j]) pairs, sorted by (col, row),
    deduplicated, column counts prefix-summed. Small inputs only."""
    rows, cols = [], []
    for sl in sl_lists:
        sl = np.asarray(sl, dtype=np.int64)
        rows.append(np.repeat(sl, len(sl)))
        cols.append(np.tile(sl, len(sl)))
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    order = np.lexsort((rows, cols))
    rows, cols = rows[order], cols[order]
    keep = np.ones(len(rows), dtype=bool)
    keep[1:] = (rows[1:] != rows[:-1]) | (cols[1:] != cols[:-1])
    rows, cols = rows[keep], cols[keep]
    indptr = np.zeros(nv + 1, dtype=np.int64)
    np.add.at(indptr, cols + 1, 1)
    indptr = np.cumsum(indptr)
    return indptr.astype(np.int32), rows.astype(np.int32)


def sparse_init(idx, offsets, nv):
    """Same canonical pattern as sparse_init_pairs (sorted unique rows per
    column), computed as the nonzero structure of B^T B with B the
    searchlight-by-vertex incidence matrix (spec: src/hyperalignment/sparse.py:29-35)."""
    import scipy.sparse as sp

    count = len(offsets) - 1
    sizes = np.diff(np.asarray(offsets, dtype=np.int64))
    rows = np.repeat(np.arange(count, dtype=np.int64), sizes)
    B = sp.csr_matrix((np.ones(len(idx), dtype=np.float32), (rows, np.asarray(idx, dtype=np.int64))),
                      shape=(count, nv))
    P = (B.T.tocsr() @ B).tocsc()
    P.sort_indices()
    return P.indptr.astype(np.int64), P.indices.astype(np.int32)


# ---------------------------------------------------------------------------
# local Procrustes
# ---------------------------------------------------------------------------

def safe_svd(A):
    """dgesdd, then dgesvd on failure (csrc/ha_searchlight.c:593-605)."""
    try:
        return sla.svd(A, full_matrices=False, lapack_driver="gesdd")
    except np.linalg.LinAlgError:
        return sla.svd(A, full_matrices=False, lapack_driver="gesvd")


def det_sign(T):
    """Sign of det(T) from LU with partial pivoting, as csrc/ha_searchlight.c:612-624."""
    lu, piv = sla.lu_factor(T, check_finite=False)
    sign = 1.0
    for i in range(T.shape[0]):
        if lu[i, i] < 0.0:
            sign = -sign
        if piv[i] != i:
            sign = -sign
    return sign


def procrustes(Xs, Ys, reflection=True, scaling=False, return_sigma=False):
    """Local Procrustes, csrc/ha_searchlight.c:578-659 (col-major workspace
    variant of csrc/ha_procrustes.c:17-127; spec src/hyperalignment/procrustes.py:61-75).

    A = Xs^T Ys; A = U s Vt; T = U Vt; if not reflection and det(T) < 0:
    s[-1] = -s[-1], T -= 2 u_last v_last^T; if scaling: T *= sum(s) /
    (sum_j var(Xs[:, j]) * n) with population variances.
    """
    A = Xs.T @ Ys
    U, s, Vt = safe_svd(A)
    T = U @ Vt
    if not reflection:
        if det_sign(T) < 0.0:
            s[-1] *= -1.0
            T = T - 2.0 * np.outer(U[:, -1], Vt[-1, :])
    if scaling:
        n = Xs.shape[0]
        mean = Xs.mean(axis=0)
        var = ((Xs - mean) ** 2).sum(axis=0) / n
        T = T * (s.sum() / (var.sum() * n))
    if return_sigma:
        return T, s
    return T


def polar_newton(A, reflection=True, scaling=False, X=None):
    """ha_polar_newton (csrc/ha_procrustes.c:271-366) at its limit: the Newton
    iteration X <- (X + X^-T)/2 from A converges to the orthogonal polar factor
    T = U Vt of A = U s Vt (nonsingular A). Without reflections, det(T) < 0
    (LU sign of T, :310-326) negates T's last column (:328-331). With
    scaling and X (M x N) given, T *= trace(T^T A) / (sum_j var(X[:, j]) * M),
    population variances (:335-362)."""
    A = np.asarray(A, dtype=np.float64)
    U, s, Vt = safe_svd(A)
    T = U @ Vt
    if not reflection and det_sign(T) < 0.0:
        T[:, -1] *= -1.0
    if scaling and X is not None:
        X = np.asarray(X, dtype=np.float64)
        M = X.shape[0]
        var = ((X - X.mean(axis=0)) ** 2).sum(axis=0) / M
        T = T * (np.trace(T.T @ A) / (var.sum() * M))
    return T


# ---------------------------------------------------------------------------
# searchlight drivers
# ---------------------------------------------------------------------------

def _lists(idx, offsets):
    return [np.asarray(idx[offsets[s]:offsets[s + 1]], dtype=np.int64) for s in range(len(offsets) - 1)]


def local_transforms(X, Y, idx, offsets, reflection=True, scaling=False, which=None, idx_y=None, n_jobs=1):
    """Local transforms T_s of the searchlights in `which` (all by default)."""
    sl = _lists(idx, offsets)
    sly = sl if idx_y is None else _lists(idx_y, offsets)
    which = range(len(sl)) if which is None else which

    def one(s):
        return procrustes(X[:, sl[s]], Y[:, sly[s]], reflection, scaling)

    if n_jobs == 1:
        return [one(s) for s in which]
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(n_jobs) as ex:
        return list(ex.map(one, list(which)))


def searchlight_procrustes_csc(X, Y, idx, offsets, indptr, indices, data0=None, weights=None,
                               reflection=True, scaling=False, idx_y=None, n_jobs=1, which_cols=None):
    """CSC accumulation as csrc/ha_searchlight.c:147-202 with ha_sparse_scatter_add
    (csrc/ha_sparse.c:128-142): data[pos(slx[i], sly[j])] += T[i, j] * w[i], in
    searchlight order (one thread), pairs outside the pattern skipped.

    which_cols: optional set of columns to compute (only searchlights touching
    them are solved); other columns keep data0.
    """
    indptr = np.asarray(indptr, dtype=np.int64)
    indices = np.asarray(indices, dtype=np.int64)
    data = np.zeros(len(indices)) if data0 is None else np.array(data0, dtype=np.float64)
    slx = _lists(idx, offsets)
    sly = slx if idx_y is None else _lists(idx_y, offsets)
    count = len(slx)
    if which_cols is None:
        todo = list(range(count))
        colmask = None
    else:
        colmask = np.zeros(int(len(indptr) - 1), dtype=bool)
        colmask[list(which_cols)] = True
        todo = [s for s in range(count) if colmask[sly[s]].any()]
    Ts = local_transforms(X, Y, idx, offsets, reflection, scaling, which=todo, idx_y=idx_y, n_jobs=n_jobs)
    _scatter_csc(data, indptr, indices, slx, sly, offsets, weights, todo, Ts, colmask)
    return data


def _scatter_csc(data, indptr, indices, slx, sly, offsets, weights, todo, Ts, colmask=None):
    """ha_sparse_scatter_add (csrc/ha_sparse.c:128-142) of each T_s in order:
    data[pos(slx[i], sly[j])] += T[i, j] * w[i]; pairs outside the pattern skipped."""
    for s, T in zip(todo, Ts):
        rows, cols = slx[s], sly[s]
        w = weights[offsets[s]:offsets[s + 1]] if weights is not None else None
        for j, c in enumerate(cols):
            if colmask is not None and not colmask[c]:
                continue
            seg = indices[indptr[c]:indptr[c + 1]]
            pos = np.searchsorted(seg, rows)
            ok = (pos < len(seg))
            ok[ok] = seg[pos[ok]] == rows[ok]
            val = T[:, j] * w if w is not None else T[:, j] * 1.0
            data[indptr[c] + pos[ok]] += val[ok]


def searchlight_procrustes_dense(X, Y, idx, offsets, dists, radius, T_out=None, reflection=True,
                                 scaling=False, n_jobs=1):
    """csrc/ha_searchlight.c:950-1106 (CPU FP64, one thread's order):
    weights as :970-1001, then T_out[sl_i, sl_j] += T[i, j] * w[i] (:1072-1085)."""
    nv = X.shape[1]
    if T_out is None:
        T_out = np.zeros((nv, nv))
    w = searchlight_weights_flat(idx, offsets, dists, radius, nv)
    sl = _lists(idx, offsets)
    Ts = local_transforms(X, Y, idx, offsets, reflection, scaling, n_jobs=n_jobs)
    for s, T in enumerate(Ts):
        ws = w[offsets[s]:offsets[s + 1]]
        T_out[np.ix_(sl[s], sl[s])] += T * ws[:, None]
    return T_out


def dense_from_csc(indptr, indices, data, n):
    out = np.zeros((n, n))
    for c in range(n):
        out[indices[indptr[c]:indptr[c + 1]], c] = data[indptr[c]:indptr[c + 1]]
    return out


# ---------------------------------------------------------------------------
# ridge
# ---------------------------------------------------------------------------

def ridge(X, Y, alpha):
    """Ridge betas through the thin SVD of X, as src/hyperalignment/ridge.py:7-12
    and csrc/ha_ridge.c:13-54: X = U diag(s) Vt, d = s / (alpha + s^2),
    betas = Vt^T (d * (U^T Y)). Y may be 1-D (csrc/ha_ridge.c:56-95)."""
    X = np.asarray(X, dtype=np.float64)
    Y = np.asarray(Y, dtype=np.float64)
    U, s, Vt = safe_svd(X)
    d = s / (alpha + s * s)
    UtY = U.T @ Y
    return Vt.T @ (d[:, None] * UtY if UtY.ndim == 2 else d * UtY)


def searchlight_ridge_csc(X, Y, idx, offsets, indptr, indices, alpha, data0=None, weights=None, idx_y=None,
                          which_cols=None):
    """csrc/ha_searchlight.c:1378-1429: ridge betas of every searchlight,
    scattered in searchlight order with ha_sparse_scatter_add (row weights).

    which_cols: optional set of columns to compute (only searchlights touching
    them are solved); other columns keep data0.
    """
    indptr = np.asarray(indptr, dtype=np.int64)
    indices = np.asarray(indices, dtype=np.int64)
    data = np.zeros(len(indices)) if data0 is None else np.array(data0, dtype=np.float64)
    slx = _lists(idx, offsets)
    sly = slx if idx_y is None else _lists(idx_y, offsets)
    colmask = None
    todo = list(range(len(slx)))
    if which_cols is not None:
        colmask = np.zeros(int(len(indptr) - 1), dtype=bool)
        colmask[list(which_cols)] = True
        todo = [s for s in todo if colmask[sly[s]].any()]
    Ts = [ridge(X[:, slx[s]], Y[:, sly[s]], alpha) for s in todo]
    _scatter_csc(data, indptr, indices, slx, sly, offsets, weights, todo, Ts, colmask)
    return data


# ---------------------------------------------------------------------------
# searchlights on a mesh
# ---------------------------------------------------------------------------

def mesh_searchlights(coords, faces, radius, centers=None):
    """Searchlights along the mesh edges, for the checker: per centre, a full
    Dijkstra (heapq) over the edge graph (edge length = Euclidean distance of
    its end points), the vertices with distance < radius, sorted by vertex id.
    The reference takes these lists from neuroboros (benchmark_neuroboros.py:153),
    absent here; parity with neuroboros itself is unpinned."""
    import heapq
    coords = np.asarray(coords, dtype=np.float64)
    nv = coords.shape[0]
    nbr = [set() for _ in range(nv)]
    for a, b, c in np.asarray(faces):
        for u, v in ((a, b), (b, c), (c, a)):
            if u != v:
                nbr[u].add(int(v))
                nbr[v].add(int(u))
    centers = range(nv) if centers is None else centers
    sls, dists = [], []
    for c0 in centers:
        dist = {int(c0): 0.0}
        done = set()
        h = [(0.0, int(c0))]
        while h:
            d, v = heapq.heappop(h)
            if v in done or d > dist[v]:
                continue
            done.add(v)
            for u in nbr[v]:
                nd = d + float(np.linalg.norm(coords[v] - coords[u]))
                if nd < radius and nd < dist.get(u, np.inf):
                    dist[u] = nd
                    heapq.heappush(h, (nd, u))
        keep = sorted(v for v in done)
        sls.append(np.array(keep, dtype=np.int64))
        dists.append(np.array([dist[v] for v in keep]))
    return sls, dists


# ---------------------------------------------------------------------------
# templates
# ---------------------------------------------------------------------------

def zscore_columns(X):
    """One-pass column z-score of csrc/ha_linalg.c:263-291 (var < 1e-30 -> 0)."""
    X = np.array(X, dtype=np.float64)
    M = X.shape[0]
    inv_M = 1.0 / M
    sums = X.sum(axis=0)
    sum2s = (X * X).sum(axis=0)
    mean = sums * inv_M
    var = sum2s * inv_M - mean * mean
    inv_std = np.where(var < 1e-30, 0.0, 1.0 / np.sqrt(np.maximum(var, 1e-300)))
    return (X - mean) * inv_std


def template_procrustes(dms, reflection=True, scaling=False, zscore_common=True, level2_iter=1):
    """csrc/ha_template.c:14-114 (spec src/hyperalignment/local_template.py:165-221)."""
    ns = len(dms)
    common = np.array(dms[0], dtype=np.float64)
    aligned = [np.array(dms[0], dtype=np.float64)]
    for s in range(1, ns):
        T = procrustes(dms[s], common, reflection, scaling)
        a = dms[s] @ T
        if zscore_common:
            a = zscore_columns(a)
        aligned.append(a)
        common = (common + a) * 0.5
        if zscore_common:
            common = zscore_columns(common)
    for _ in range(level2_iter):
        common = np.zeros_like(common)
        for s in range(ns):
            common += aligned[s]
        for s in range(ns):
            ref = (common - aligned[s]) * (1.0 / (ns - 1))
            if zscore_common:
                ref = zscore_columns(ref)
            T = procrustes(dms[s], ref, reflection, scaling)
            aligned[s] = dms[s] @ T
    tpl = np.zeros_like(common)
    for s in range(ns):
        tpl += aligned[s]
    if zscore_common:
        tpl = zscore_columns(tpl)
    else:
        tpl *= 1.0 / ns
    return tpl


def searchlight_template(dms, idx, offsets, dists, radius, reflection=True, scaling=False,
                         zscore_common=True, level2_iter=2, which=None, n_jobs=1):
    """Searchlight template: local templates (csrc/ha_template.c:14-114 on the
    searchlight's columns) combined with per-vertex weights as
    src/hyperalignment/searchlight.py:215-236 (tpl[:, sl] += local * w), added
    in searchlight order (n_jobs only spreads the local templates over threads)."""
    nv = dms[0].shape[1]
    w = searchlight_weights_flat(idx, offsets, dists, radius, nv)
    tpl = np.zeros_like(np.asarray(dms[0], dtype=np.float64))
    sl = _lists(idx, offsets)
    which = list(range(len(sl)) if which is None else which)

    def one(s):
        return template_procrustes([d[:, sl[s]] for d in dms], reflection, scaling, zscore_common, level2_iter)

    if n_jobs == 1:
        locs = [one(s) for s in which]
    else:
        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(n_jobs) as ex:
            locs = list(ex.map(one, which))
    for s, loc in zip(which, locs):
        tpl[:, sl[s]] += loc * w[offsets[s]:offsets[s + 1]][None, :]
    return tpl


# ---------------------------------------------------------------------------
# applying and scoring a transform
# ---------------------------------------------------------------------------

def apply_transform(X, indptr, indices, data, ncols):
    """Yhat = X @ W for the CSC transform W (rows: source vertices), as
    benchmark_neuroboros.py:229-232 (X_test @ Ws). Column c accumulates
    X[:, indices[q]] * data[q] over its nonzeros in index order, each multiply
    and add rounded separately (the order of a sequential CSC product)."""
    X = np.asarray(X, dtype=np.float64)
    Y = np.zeros((X.shape[0], ncols))
    for c in range(ncols):
        acc = np.zeros(X.shape[0])
        for q in range(indptr[c], indptr[c + 1]):
            acc = acc + X[:, indices[q]] * data[q]
        Y[:, c] = acc
    return Y


def vertex_correlation(A, B):
    """np.mean(zscore(A, axis=0) * zscore(B, axis=0), axis=0) with population
    z-scores (scipy.stats.zscore, ddof=0), as benchmark_neuroboros.py:204-211."""
    A = np.asarray(A, dtype=np.float64)
    B = np.asarray(B, dtype=np.float64)
    with np.errstate(invalid="ignore", divide="ignore"):
        za = (A - A.mean(axis=0)) / A.std(axis=0)
        zb = (B - B.mean(axis=0)) / B.std(axis=0)
        return (za * zb).mean(axis=0)
