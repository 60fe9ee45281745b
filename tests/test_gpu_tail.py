"""The uneven-size variant of the full-hemisphere workload ("c2tail": sizes
37-530 with a long tail, every iteration-tile bucket populated, 168
searchlights with k >= n): one default pass per method, checked against the
FP64 oracle on sampled columns that no rank-deficient searchlight touches
(there the polar factor is not unique), and on rank-deficient searchlights by
the invariants of an optimal orthogonal Procrustes solution. The polar
iteration solves them all: no FP64 fallback.

Tolerance: max |T - T_fp64| <= 1e-4 on the sampled columns, the bar of the
other polar tests, and 1e-4 on the invariants.
"""

import numpy as np
import pytest

import hyperalignment_amd as ha
from hyperalignment_amd import learn, synth
from oracle import ha_oracle as O

pytestmark = pytest.mark.gpu

TOL_T = 1e-4


@pytest.fixture(scope="module")
def tail():
    ha.init()
    w = synth.make_workload("c2tail")
    print(f"[tail] k {w.sls.sizes.min()}-{w.sls.sizes.max()}", flush=True)
    yield w
    ha.set_model(None, None)
    ha.set_options(method="auto", tile=0)


@pytest.mark.parametrize("method", ["polar", "polar_learned"])
def test_tail_pass(tail, method):
    w = tail
    s = w.sls
    nt = w.X.shape[0]
    if method == "polar_learned":
        learn.install()
    ha.set_options(method=method, tile=0)
    k = s.sizes
    deficient = np.flatnonzero(k >= nt)
    assert len(deficient) > 0
    picks = deficient[[0, len(deficient) // 2, len(deficient) - 1]]
    plan = ha.Plan(nt, s.nv, s.indices, s.offsets, s.dists, s.radius)
    try:
        plan.load(w.X, w.Y)
        plan.execute()
        indptr, indices, data = plan.csc()
        st = plan.stats()
        steps, flags = plan.steps()
        tiles = plan.tiles()
        local = plan.local(picks)
    finally:
        plan.close()
    print(f"[tail] {method}: {st['ms_total']:.1f} ms, tiles {tiles.tolist()}, "
          f"fallback {st['n_fallback']}, steps of k >= n {steps[deficient].min()}-{steps[deficient].max()}",
          flush=True)
    assert st["n_unconverged"] == 0
    assert st["n_fallback"] == 0
    for sidx, T in zip(picks, local):
        l = s.indices[s.offsets[sidx]:s.offsets[sidx + 1]]
        Xs, Ys = w.X[:, l], w.Y[:, l]
        T = T.astype(np.float64)
        assert np.abs(T.T @ T - np.eye(len(l))).max() < TOL_T
        H = T.T @ (Xs.T @ Ys)
        assert np.abs(H - H.T).max() < TOL_T * np.abs(H).max()
        r, ro = np.linalg.norm(Xs @ T - Ys), np.linalg.norm(Xs @ O.procrustes(Xs, Ys) - Ys)
        assert abs(r - ro) <= TOL_T * ro
    touched = np.unique(np.concatenate([s.indices[s.offsets[i]:s.offsets[i + 1]] for i in deficient]))
    free = np.setdiff1d(np.arange(s.nv), touched)
    cols = [int(free[j]) for j in (0, len(free) // 3, 2 * len(free) // 3, len(free) - 1)]
    wts = O.searchlight_weights_flat(s.indices, s.offsets, s.dists, s.radius, s.nv)
    ref = O.searchlight_procrustes_csc(w.X, w.Y, s.indices, s.offsets, indptr, indices, weights=wts,
                                       which_cols=cols, n_jobs=8)
    for c in cols:
        sl = slice(indptr[c], indptr[c + 1])
        assert np.abs(data[sl] - ref[sl]).max() < TOL_T, c
