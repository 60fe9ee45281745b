"""GPU parity of the HIP backend against the CPU FP64 oracle, through the C-ABI.

Tolerances: integer work (pattern, index lists) and the FP64 weights are
bit-exact; the FP64 Jacobi path is within 1e-6 of LAPACK (its local transforms
are stored in FP32); FP32 / bf16x3 paths are within max-abs 1e-4 of the
oracle on T, the bar the reference's GPU target states.
"""

import numpy as np
import scipy.sparse
import pytest

import hyperalignment_amd as ha
from hyperalignment_amd import synth
from oracle import ha_oracle as O
from tests import refcases

pytestmark = pytest.mark.gpu

TOL_T = 1e-4     # max |T - T_fp64| for FP32 / bf16x3 local transforms
TOL_64 = 1e-6    # FP64 Jacobi (FP32 storage of local transforms)


@pytest.fixture(scope="module", autouse=True)
def hip():
    ha.init()
    yield
    ha.set_options(method="auto")


@pytest.fixture(scope="module")
def c0():
    w = synth.make_workload("c0")
    s = w.sls
    T = O.searchlight_procrustes_dense(w.X, w.Y, s.indices, s.offsets, s.dists, s.radius)
    return w, T


def run_dense(w, method, reflection=True, scaling=False):
    ha.set_options(method=method)
    s = w.sls
    return ha.searchlight_procrustes_flat(w.X, w.Y, s.indices, s.offsets, s.dists, s.radius,
                                          reflection=reflection, scaling=scaling)


# ---- the reference's known-answer tests, on the GPU ----

def test_ref_procrustes_identity():
    X = refcases.identity_case()
    T = ha.procrustes(X, X)
    assert np.abs(np.abs(T) - np.eye(20)).max() < 1e-8


def test_ref_procrustes_known_rotation():
    X, Y, R = refcases.rotation_case()
    T = ha.procrustes(X, Y)
    assert np.abs(X @ T - Y).max() < 1e-8


def test_ref_procrustes_no_reflection():
    X, Y = refcases.no_reflection_case()
    T = ha.procrustes(X, Y, reflection=False)
    assert np.linalg.det(T) > 0
    assert np.abs(T - O.procrustes(X, Y, False, False)).max() < 1e-8


def test_ref_procrustes_scaling():
    rng = np.random.default_rng(3)
    X = rng.standard_normal((80, 9))
    Y = rng.standard_normal((80, 9))
    assert np.abs(ha.procrustes(X, Y, scaling=True) - O.procrustes(X, Y, True, True)).max() < 1e-8


def test_ref_sparse_init_toy():
    m = ha.sparse_init(refcases.SPARSE_TOY, 5)
    assert m.nnz == 21 and m.shape == (5, 5)
    p, i = O.sparse_init_pairs(refcases.SPARSE_TOY, 5)
    assert np.array_equal(m.indptr, p) and np.array_equal(m.indices, i)
    assert np.all(m.data == 0)


def test_ref_scatter_add_identity():
    m = ha.sparse_init([[0, 1, 2]], 3)
    ha.sparse_scatter_add(m, [0, 1, 2], [0, 1, 2], np.eye(3))
    assert np.array_equal(m.toarray(), np.eye(3))


def test_ref_weights_uniform_toy():
    w = ha.searchlight_weights(refcases.WEIGHTS_TOY)
    sums = np.zeros(6)
    for l, x in zip(refcases.WEIGHTS_TOY, w):
        np.add.at(sums, l, x)
    assert np.abs(sums - 1).max() < 1e-12


# ---- bit-exact integer and FP64 work ----

@pytest.mark.parametrize("uniform", [True, False])
def test_weights_bitexact_c0(uniform):
    w = synth.make_workload("c0")
    s = w.sls
    d = None if uniform else s.dist_lists()
    got = np.concatenate(ha.searchlight_weights(s.lists(), d, s.radius))
    ref = O.searchlight_weights_flat(s.indices, s.offsets, None if uniform else s.dists, s.radius, s.nv)
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))


@pytest.mark.parametrize("name", ["c0", "c1"])
def test_pattern_bitexact(name):
    w = synth.make_workload(name)
    s = w.sls
    m = ha.sparse_init(s.lists(), s.nv)
    p, i = O.sparse_init(s.indices, s.offsets, s.nv)
    assert np.array_equal(m.indptr.astype(np.int64), p)
    assert np.array_equal(m.indices, i)


# ---- dense entry point vs the oracle ----

@pytest.mark.parametrize("method,tol", [("jacobi64", TOL_64), ("jacobi32", TOL_T), ("polar", TOL_T)])
def test_dense_c0(c0, method, tol):
    w, Tref = c0
    T = run_dense(w, method)
    err = np.abs(T - Tref).max()
    assert err < tol, (method, err)
    assert np.array_equal(T != 0, Tref != 0)


@pytest.mark.parametrize("method", ["jacobi32", "polar"])
def test_identity_invariant(method):
    w = synth.make_workload("c0")
    s = w.sls
    ha.set_options(method=method)
    T = ha.searchlight_procrustes_flat(w.X, w.X, s.indices, s.offsets, s.dists, s.radius)
    assert np.abs(T - np.eye(s.nv)).max() < 1e-4


def test_dense_accumulates(c0):
    w, Tref = c0
    s = w.sls
    ha.set_options(method="jacobi64")
    T0 = np.full((s.nv, s.nv), 0.25)
    T = ha.searchlight_procrustes_flat(w.X, w.Y, s.indices, s.offsets, s.dists, s.radius, T=T0.copy())
    ref = O.searchlight_procrustes_dense(w.X, w.Y, s.indices, s.offsets, s.dists, s.radius, T_out=T0.copy())
    assert np.abs(T - ref).max() < TOL_64


@pytest.mark.parametrize("method,tol", [("polar", TOL_T), ("jacobi64", TOL_64)])
@pytest.mark.parametrize("scaling", [False, True])
def test_reflection_false_c0(method, tol, scaling):
    """isReflection = false: the polar path fixes det(Q) < 0 after the iteration
    (LU sign of Q, smallest right singular vector of the FP64 A, rank-1 flip),
    the FP64 path inside its Jacobi SVD; both against LAPACK's dgesdd + LU sign."""
    w = synth.make_workload("c0")
    s = w.sls
    ha.set_options(method=method)
    T = ha.searchlight_procrustes_flat(w.X, w.Y, s.indices, s.offsets, s.dists, s.radius, reflection=False,
                                       scaling=scaling)
    ref = O.searchlight_procrustes_dense(w.X, w.Y, s.indices, s.offsets, s.dists, s.radius, reflection=False,
                                         scaling=scaling)
    assert np.abs(T - ref).max() < tol


def test_reflection_false_c1_local():
    """The polar path's reflection fix at kp 96-128: every local transform has
    det = +1 and matches the FP64 path's to the polar bar."""
    w = synth.make_workload("c1")
    which = np.arange(0, w.sls.count, 61)[:96]
    from hyperalignment_amd import learn
    idx, off, d = learn.sub_searchlights(w.sls, which)
    out = {}
    for method in ("polar", "jacobi64"):
        plan = ha.Plan(w.X.shape[0], w.sls.nv, idx, off, d, w.sls.radius)
        plan.load(w.X, w.Y)
        ha.set_options(method=method)
        plan.execute(reflection=False)
        out[method] = plan.local(np.arange(len(which)))
        st, fl = plan.steps()
        if method == "polar":
            assert not fl.any()
        plan.close()
    for Tp, Tj in zip(out["polar"], out["jacobi64"]):
        dp = np.linalg.det(Tp.astype(np.float64))
        assert abs(dp - 1.0) < 1e-3
        assert np.abs(Tp - Tj).max() < TOL_T
    print(f"[refl] {len(which)} searchlights checked", flush=True)


@pytest.mark.parametrize("method", ["jacobi32", "polar"])
def test_scaling_c0(method):
    w = synth.make_workload("c0")
    s = w.sls
    T = run_dense(w, method, scaling=True)
    ref = O.searchlight_procrustes_dense(w.X, w.Y, s.indices, s.offsets, s.dists, s.radius, scaling=True)
    assert np.abs(T - ref).max() < TOL_T


def test_deterministic(c0):
    w, _ = c0
    a = run_dense(w, "polar")
    b = run_dense(w, "polar")
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


# ---- CSC entry point ----

def test_csc_api_c0(c0):
    w, _ = c0
    s = w.sls
    ha.set_options(method="jacobi32")
    m = ha.sparse_init(s.lists(), s.nv)
    wts = O.searchlight_weights_flat(s.indices, s.offsets, s.dists, s.radius, s.nv)
    wl = [wts[s.offsets[i]:s.offsets[i + 1]] for i in range(s.count)]
    m.data[:] = 0.5
    ha.searchlight_procrustes_csc(w.X.T.copy().T, w.Y, s.lists(), m, weights=wl)
    ref = O.searchlight_procrustes_csc(w.X, w.Y, s.indices, s.offsets, m.indptr, m.indices,
                                       data0=np.full(m.nnz, 0.5), weights=wts)
    assert np.abs(m.data - ref).max() < TOL_T


def test_csc_api_unweighted_separate_y_lists():
    rng = np.random.default_rng(7)
    nv, nt = 30, 40
    X = rng.standard_normal((nt, nv))
    Y = rng.standard_normal((nt, nv))
    lx = [np.sort(rng.choice(nv, 6, replace=False)) for _ in range(12)]
    ly = [np.sort(rng.choice(nv, 6, replace=False)) for _ in range(12)]
    m = ha.sparse_init(lx, nv)
    ha.set_options(method="jacobi64")
    ha.searchlight_procrustes_csc(X, Y, lx, m, sls_Y=ly)
    sx = synth.toy_searchlights(lx, nv=nv)
    sy = synth.toy_searchlights(ly, nv=nv)
    ref = O.searchlight_procrustes_csc(X, Y, sx.indices, sx.offsets, m.indptr, m.indices, idx_y=sy.indices)
    assert np.abs(m.data - ref).max() < TOL_64


def test_csc_long_columns_global_path():
    """A caller pattern with five full columns (3,300 rows, longer than the 3,072
    the scatter holds in LDS): those columns accumulate in global memory, the
    others in LDS, in one pass; rows outside every searchlight stay zero."""
    rng = np.random.default_rng(5)
    nv, nt = 3300, 40
    X = rng.standard_normal((nt, nv))
    Y = rng.standard_normal((nt, nv))
    lx = [np.sort(rng.choice(nv, 12, replace=False)) for _ in range(60)]
    base = ha.sparse_init(lx, nv)
    full = {int(lx[i][0]) for i in range(5)}
    cols, rows = [], []
    for c in range(nv):
        r = np.arange(nv) if c in full else base.indices[base.indptr[c]:base.indptr[c + 1]]
        rows.append(r)
        cols.append(np.full(len(r), c))
    rows, cols = np.concatenate(rows), np.concatenate(cols)
    m = scipy.sparse.csc_matrix((np.zeros(len(rows)), (rows, cols)), shape=(nv, nv))
    m.sort_indices()
    assert np.diff(m.indptr).max() == nv
    ha.set_options(method="jacobi64")
    ha.searchlight_procrustes_csc(X, Y, lx, m)
    sx = synth.toy_searchlights(lx, nv=nv)
    ref = O.searchlight_procrustes_csc(X, Y, sx.indices, sx.offsets, m.indptr, m.indices)
    assert np.abs(m.data - ref).max() < TOL_64
    for c in full:
        assert np.count_nonzero(m.data[m.indptr[c]:m.indptr[c + 1]]) > 0


# ---- larger config: sampled columns ----

def test_c1_polar_columns():
    w = synth.make_workload("c1")
    s = w.sls
    plan = ha.Plan(w.X.shape[0], s.nv, s.indices, s.offsets, s.dists, s.radius)
    plan.load(w.X, w.Y)
    ha.set_options(method="polar")
    plan.execute()
    indptr, indices, data = plan.csc()
    p, i = O.sparse_init(s.indices, s.offsets, s.nv)
    assert np.array_equal(indptr.astype(np.int64), p) and np.array_equal(indices, i)
    cols = [0, 17, 5000, 10241]
    wts = O.searchlight_weights_flat(s.indices, s.offsets, s.dists, s.radius, s.nv)
    ref = O.searchlight_procrustes_csc(w.X, w.Y, s.indices, s.offsets, indptr, indices, weights=wts,
                                       which_cols=cols, n_jobs=8)
    for c in cols:
        sl = slice(indptr[c], indptr[c + 1])
        assert np.abs(data[sl] - ref[sl]).max() < TOL_T
    st = plan.stats()
    assert st["n_unconverged"] == 0
    plan.close()


@pytest.mark.parametrize("name", ["c0", "c1"])
def test_learned_method_outside_its_training_sizes(name):
    """The learned scaling was trained on searchlights of k 357-441; on smaller
    ones (k 40-111) its prediction is held to what A proves about sigma_max, so
    every searchlight still runs the polar iteration (none is left to the FP64
    path) and the columns match the oracle."""
    from hyperalignment_amd import learn
    w = synth.make_workload(name)
    s = w.sls
    learn.install(learn.MODEL_PATH)
    plan = ha.Plan(w.X.shape[0], s.nv, s.indices, s.offsets, s.dists, s.radius)
    try:
        plan.load(w.X, w.Y)
        ha.set_options(method="polar_learned")
        plan.execute()
        st = plan.stats()
        print(f"[learned {name}] fallback {st['n_fallback']}, orth flagged {st['n_orth_flagged']}, "
              f"mean steps {st['steps_total'] / s.count:.2f}, {st['ms_total']:.1f} ms", flush=True)
        assert st["n_fallback"] == 0 and st["n_orth_flagged"] == 0 and st["n_unconverged"] == 0
        indptr, indices, data = plan.csc()
        cols = [0, 17, s.nv // 2, s.nv - 1]
        wts = O.searchlight_weights_flat(s.indices, s.offsets, s.dists, s.radius, s.nv)
        ref = O.searchlight_procrustes_csc(w.X, w.Y, s.indices, s.offsets, indptr, indices, weights=wts,
                                           which_cols=cols, n_jobs=8)
        for c in cols:
            sl = slice(indptr[c], indptr[c + 1])
            assert np.abs(data[sl] - ref[sl]).max() < TOL_T
    finally:
        ha.set_options(method="polar")
        plan.close()


# ---- rank deficiency: parity by invariants (the polar factor is not unique) ----

def _check_rank_deficient(X, Y, T, tol):
    k = X.shape[1]
    assert np.abs(T.T @ T - np.eye(k)).max() < tol
    H = T.T @ (X.T @ Y)
    assert np.abs(H - H.T).max() < tol * max(1.0, np.abs(H).max())
    assert np.linalg.eigvalsh(0.5 * (H + H.T)).min() > -tol * max(1.0, np.abs(H).max())
    To = O.procrustes(X, Y)
    r, ro = np.linalg.norm(X @ T - Y), np.linalg.norm(X @ To - Y)
    assert abs(r - ro) <= tol * max(1.0, ro)


def test_rank_deficient_local():
    rng = np.random.default_rng(11)
    X = rng.standard_normal((8, 20))
    Y = rng.standard_normal((8, 20))
    _check_rank_deficient(X, Y, ha.procrustes(X, Y), 1e-9)


@pytest.mark.parametrize("method,tol", [("polar", TOL_T), ("jacobi32", 2e-6)])
def test_rank_deficient_searchlights(method, tol):
    """k >= n_samples: the polar path iterates from the smallest bucket's
    schedule, which takes the rounding-noise singular values to 1 as well (one
    orthogonal completion), and the FP32 Jacobi path hands them to the FP64
    fallback. Every local transform is an optimal orthogonal Procrustes
    solution, to the accuracy of the path that made it."""
    rng = np.random.default_rng(12)
    nv, nt = 60, 10
    X = rng.standard_normal((nt, nv))
    Y = rng.standard_normal((nt, nv))
    lists = [np.sort(rng.choice(nv, 16, replace=False)) for _ in range(20)]
    sl = synth.toy_searchlights(lists, nv=nv)
    plan = ha.Plan(nt, nv, sl.indices, sl.offsets, None, 0.0)
    plan.load(X, Y)
    ha.set_options(method=method)
    plan.execute()
    st, fl = plan.steps()
    print(f"[rank] {method}: fallback {int(fl.sum())} of {len(lists)}, steps {st.tolist()}", flush=True)
    assert fl.all() if method == "jacobi32" else not fl.any()
    for s, T in enumerate(plan.local(np.arange(len(lists)))):
        l = lists[s]
        _check_rank_deficient(X[:, l], Y[:, l], T.astype(np.float64), tol)
    plan.close()


# ---- captured passes and the tile choice ----

def test_graph_replay_matches_eager():
    """A pass captured in a hipGraph and replayed gives the eager result bit for
    bit on every replay, with eager passes and host reads between replays, and
    the stage events recorded inside the graph time every replay."""
    w = synth.make_workload("c1")
    s = w.sls
    plan = ha.Plan(w.X.shape[0], s.nv, s.indices, s.offsets, s.dists, s.radius)
    plan.load(w.X, w.Y)
    ha.set_options(method="polar", use_graph=0)
    plan.execute()
    ref = plan.csc()[2].copy()
    try:
        for rep in range(3):
            ha.set_options(use_graph=1)
            plan.execute()
            st = plan.stats()
            print(f"[graph] replay {rep}: {st['ms_total']:.2f} ms, iteration {st['kernel_ms_polar_avg']:.2f} ms",
                  flush=True)
            assert st["graph_replay"] == 1
            assert st["ms_total"] > 0 and st["kernel_ms_polar_avg"] > 0
            assert st["ms_crosscov"] > 0 and st["ms_scatter"] > 0
            got = plan.csc()[2]
            assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), rep
            ha.set_options(use_graph=0)
            plan.execute()
            assert plan.stats()["graph_replay"] == 0
            assert np.array_equal(plan.csc()[2].view(np.uint64), ref.view(np.uint64))
    finally:
        ha.set_options(use_graph=1)
        plan.close()


def test_default_passes_replay_after_the_first():
    """By default the first pass of a plan runs eagerly and the later ones replay
    the captured graph; a change of options recaptures, and the results stay
    identical to the eager ones throughout."""
    w = synth.make_workload("c1")
    s = w.sls
    plan = ha.Plan(w.X.shape[0], s.nv, s.indices, s.offsets, s.dists, s.radius)
    plan.load(w.X, w.Y)
    try:
        ha.set_options(method="polar", use_graph=1)
        plan.execute()
        assert plan.stats()["graph_replay"] == 0
        ref = plan.csc()[2].copy()
        for rep in range(2):
            plan.execute()
            assert plan.stats()["graph_replay"] == 1
            assert np.array_equal(plan.csc()[2].view(np.uint64), ref.view(np.uint64)), rep
        # scaling is part of the captured pass: a new capture, then replays of it
        plan.execute(scaling=True)
        sc = plan.csc()[2].copy()
        assert not np.array_equal(sc, ref)
        plan.execute(scaling=True)
        assert np.array_equal(plan.csc()[2].view(np.uint64), sc.view(np.uint64))
        ha.set_options(use_graph=0)
        plan.execute(scaling=True)
        assert np.array_equal(plan.csc()[2].view(np.uint64), sc.view(np.uint64))
    finally:
        ha.set_options(use_graph=1)
        plan.close()


def test_graph_replay_without_reflections():
    """The pass without reflections (the LU-sign and reflection-fix kernels,
    their queues and the det sign buffer) captured and replayed gives the
    eager result bit for bit, also when passes with reflections come between
    (each change of the flag captures anew)."""
    w = synth.make_workload("c1")
    s = w.sls
    plan = ha.Plan(w.X.shape[0], s.nv, s.indices, s.offsets, s.dists, s.radius)
    plan.load(w.X, w.Y)
    try:
        ha.set_options(method="polar", use_graph=0)
        plan.execute(reflection=False)
        ref_f = plan.csc()[2].copy()
        plan.execute(reflection=True)
        ref_t = plan.csc()[2].copy()
        assert not np.array_equal(ref_f, ref_t)
        ha.set_options(use_graph=1)
        for refl, ref in ((False, ref_f), (False, ref_f), (True, ref_t), (False, ref_f), (False, ref_f)):
            plan.execute(reflection=refl)
            got = plan.csc()[2]
            print(f"[graph refl] reflection={refl} replay={plan.stats()['graph_replay']}", flush=True)
            assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), refl
    finally:
        ha.set_options(use_graph=1)
        plan.close()


def test_autotuner_report():
    """The tuner's report per kp bucket: every bucket the workload fills gets a
    launch, timed buckets keep the fastest of the three candidates, and the
    lists cover every searchlight."""
    w = synth.make_workload("c1")
    s = w.sls
    plan = ha.Plan(w.X.shape[0], s.nv, s.indices, s.offsets, s.dists, s.radius)
    plan.load(w.X, w.Y)
    try:
        ha.set_options(method="polar", tile=0, team=1)
        tn = plan.tuning()
        print(tn, flush=True)
        assert sum(b["searchlights"] for b in tn) == s.count
        for b in tn:
            assert b["tile"] in (64, 128)
            assert not b["teams"] or b["tile"] == 128
            if b["sample"]:
                ms = b["ms"]
                cand = {(64, False): ms["64"], (128, False): ms["128"], (128, True): ms["128_teams"]}
                chosen = cand[(b["tile"], b["teams"])]
                assert chosen == min(v for v in cand.values() if v)
        assert list(plan.tiles()) == [b["tile"] for b in tn]
    finally:
        plan.close()


@pytest.mark.parametrize("tile", [64, 128])
def test_forced_tile_matches_default(tile):
    """Every iteration tile gives the same transforms within the bf16x3 bar."""
    w = synth.make_workload("c1")
    s = w.sls
    plan = ha.Plan(w.X.shape[0], s.nv, s.indices, s.offsets, s.dists, s.radius)
    plan.load(w.X, w.Y)
    try:
        ha.set_options(method="polar", tile=0)
        plan.execute()
        ref = plan.csc()[2].copy()
        ha.set_options(tile=tile)
        plan.execute()
        assert set(plan.tiles().tolist()) == {tile}
        assert np.abs(plan.csc()[2] - ref).max() < TOL_T
    finally:
        ha.set_options(tile=0)
        plan.close()


# ---- the Python package's API names and the scipy .npz cache ----

def test_initialize_sparse_matrix_npz_cache(tmp_path):
    """The pattern saved through the cache is a scipy .npz; loading it back, or
    building it again, gives the reference's union pattern (sparse.py:29-42)."""
    import scipy.sparse as sp
    w = synth.make_workload("c0")
    s = w.sls
    fn = str(tmp_path / "cache" / "pattern.npz")
    m = ha.initialize_sparse_matrix(s.lists(), s.nv, cache_fn=fn)
    p, i = O.sparse_init(s.indices, s.offsets, s.nv)
    assert np.array_equal(m.indptr.astype(np.int64), p) and np.array_equal(m.indices, i)
    assert np.all(m.data == 0) and m.data.dtype == np.float64
    loaded = sp.load_npz(fn)
    assert loaded.format == "csc" and (loaded != m).nnz == 0
    again = ha.initialize_sparse_matrix(s.lists(), s.nv, cache_fn=fn)
    assert np.array_equal(again.indptr, m.indptr) and np.array_equal(again.indices, m.indices)


def test_compute_searchlight_weights_sparse():
    w = synth.make_workload("c0")
    s = w.sls
    dl = s.dist_lists()
    lists = ha.compute_searchlight_weights(s.lists(), dl, s.radius)
    mat = ha.compute_searchlight_weights(s.lists(), dl, s.radius, return_sparse=True)
    ref = O.searchlight_weights_flat(s.indices, s.offsets, s.dists, s.radius, s.nv)
    assert np.array_equal(np.concatenate(lists), ref)
    assert mat.shape == (s.count, s.nv)
    assert np.abs(np.asarray(mat.sum(axis=0)).ravel() - 1).max() < 1e-12


# ---- hand-derived driver case (row weights, non-symmetric local transforms) ----

@pytest.mark.parametrize("method,tol", [("jacobi64", TOL_64), ("jacobi32", TOL_T), ("polar", TOL_T)])
def test_driver_row_weights_kat(method, tol):
    """The known answer of tests/test_oracle_kat.py (derived from
    csrc/ha_searchlight.c:970-1001,1072-1085), on the GPU dense and CSC entries."""
    from tests import test_oracle_kat as K
    ha.set_options(method=method)
    T = ha.searchlight_procrustes_flat(K.KAT_X, K.KAT_Y, K.KAT_IDX, K.KAT_OFF, K.KAT_DIST, K.KAT_RADIUS)
    assert np.abs(T - K.KAT_T).max() < tol
    lists = [K.KAT_IDX[0:2], K.KAT_IDX[2:4]]
    m = ha.sparse_init(lists, 3)
    wts = O.searchlight_weights_flat(K.KAT_IDX, K.KAT_OFF, K.KAT_DIST, K.KAT_RADIUS, 3)
    ha.searchlight_procrustes_csc(K.KAT_X, K.KAT_Y, lists, m, weights=[wts[0:2], wts[2:4]])
    assert np.abs(m.toarray() - K.KAT_T).max() < tol


def test_unrepaired_flags_are_an_error_without_fp64_fallback():
    """With the FP64 fallback off, a pass that leaves searchlights flagged (here
    every one: a step budget of 2 stops the iteration early) returns
    kHaErrorSvd instead of scattering unfinished transforms as a success."""
    w = synth.make_workload("c0")
    s = w.sls
    plan = ha.Plan(w.X.shape[0], s.nv, s.indices, s.offsets, s.dists, s.radius)
    try:
        plan.load(w.X, w.Y)
        ha.set_options(method="polar", fp64_fallback=0, max_polar_steps=2)
        with pytest.raises(ha.HaError) as ei:
            plan.execute()
        assert ei.value.rc == -2
        ha.set_options(fp64_fallback=1)
        plan.execute()   # the same budget with the fallback on: repaired by the FP64 path
        assert plan.stats()["n_fallback"] == s.count
    finally:
        ha.set_options(method="auto", fp64_fallback=1, max_polar_steps=28)
        plan.close()


# ---- ha_polar_newton (csrc/ha_procrustes.c:271-366) ----

@pytest.mark.parametrize("n", [8, 50, 200])
def test_polar_newton_matches_oracle(n):
    """The polar factor of a square FP32 matrix against the oracle's (the limit
    of the reference's Newton iteration, pinned in tests/test_oracle_kat.py), with
    the reference's reflection rule (det < 0: last column negated)."""
    rng = np.random.default_rng(20 + n)
    A = rng.standard_normal((n, n)).astype(np.float32)
    if np.linalg.det(A.astype(np.float64)) > 0:
        A[:, 0] *= -1.0          # det < 0, so the reflection rule acts
    A64 = A.astype(np.float64)
    T = ha.polar_newton(A)
    assert np.abs(T - O.polar_newton(A64)).max() < TOL_T
    Tn = ha.polar_newton(A, reflection=False)
    assert np.linalg.det(Tn.astype(np.float64)) > 0
    assert np.abs(Tn - O.polar_newton(A64, reflection=False)).max() < TOL_T


def test_polar_newton_scaling():
    rng = np.random.default_rng(31)
    X = rng.standard_normal((60, 12))
    R = np.linalg.qr(rng.standard_normal((12, 12)))[0]
    A = (X.T @ (1.5 * X @ R)).astype(np.float32)
    T = ha.polar_newton(A, scaling=True, X=X)
    ref = O.polar_newton(A.astype(np.float64), scaling=True, X=X.astype(np.float32).astype(np.float64))
    assert np.abs(T - ref).max() < TOL_T * max(1.0, np.abs(ref).max())
