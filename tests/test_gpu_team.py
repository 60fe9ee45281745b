"""Teams of four workgroups per searchlight (128 x 128 iteration tiles): parity
with the one-workgroup kernel and the FP64 oracle, determinism, and the
give-back path when a team barrier gives up (forced from the host with a wait
budget of 1 microsecond, so team waits time out at once and the one-workgroup
kernel finishes the pass; the pass counts the teams that gave up and the
searchlights handed back). The FP32 team instance (rank-deficient
searchlights, 64 x 64 tiles) is checked at kp 32, one tile per product, by the
invariants of an optimal solution.

Tolerance: max |T - T_fp64| <= 1e-4 on sampled CSC columns, the bar of the
other polar tests; team and one-workgroup results agree to 1e-4 (their
residual sums differ in order, which can move a stopping decision by a step).
"""

import numpy as np
import pytest

import hyperalignment_amd as ha
from hyperalignment_amd import synth
from oracle import ha_oracle as O

pytestmark = pytest.mark.gpu

TOL_T = 1e-4
TOL_BF16X3_SMALL = 5e-4  # bf16x3 on 13-vertex searchlights of 16 samples (k / (n - 1) = 0.87)


def k_of(s, i):
    return int(s.offsets[i + 1] - s.offsets[i])


@pytest.fixture(scope="module", autouse=True)
def hip():
    ha.init()
    yield
    ha.set_options(method="auto", tile=0, team=1, team_spin=0)


def run(w, team, team_spin=0, method="polar"):
    ha.set_options(method=method, tile=128, team=team, team_spin=team_spin)
    s = w.sls
    plan = ha.Plan(w.X.shape[0], s.nv, s.indices, s.offsets, s.dists, s.radius)
    plan.load(w.X, w.Y)
    plan.execute()
    indptr, indices, data = plan.csc()
    st = plan.stats()
    steps, flags = plan.steps()
    plan.close()
    return indptr, indices, data, st, steps


# kp128: one 128 tile per product, so member 0 forms every product and the
# others only read what it wrote (the barrier's acquire side carries all of it)
@pytest.fixture(scope="module", params=[("c1", None, None), ("c1", 28.0, None), ("c0", 75.0, 500)],
                ids=["kp128", "kp224", "kp352"])
def case(request):
    name, radius, n = request.param
    w = synth.make_workload(name, radius=radius, n=n)
    print(f"[team] {name} radius {radius}: k {w.sls.sizes.min()}-{w.sls.sizes.max()}", flush=True)
    return w


def oracle_cols(w, indptr, indices, cols):
    s = w.sls
    wts = O.searchlight_weights_flat(s.indices, s.offsets, s.dists, s.radius, s.nv)
    return O.searchlight_procrustes_csc(np.asarray(w.X, np.float64), np.asarray(w.Y, np.float64), s.indices,
                                        s.offsets, indptr, indices, weights=wts, which_cols=cols, n_jobs=8)


def test_team_matches_solo_and_oracle(case):
    w = case
    ip, ix, d_team, st, steps = run(w, team=1)
    _, _, d_solo, st0, _ = run(w, team=0)
    print(f"[team] pass team {st['ms_total']:.1f} ms, solo {st0['ms_total']:.1f} ms", flush=True)
    assert st["n_unconverged"] == 0 and np.all(steps != 0)
    assert st["n_team_giveup"] == 0 and st["n_team_handback"] == 0
    assert np.abs(d_team - d_solo).max() < TOL_T
    cols = [0, w.sls.nv // 3, w.sls.nv - 1]
    ref = oracle_cols(w, ip, ix, cols)
    for c in cols:
        sl = slice(ip[c], ip[c + 1])
        assert np.abs(d_team[sl] - ref[sl]).max() < TOL_T


def test_team_deterministic(case):
    w = case
    a = run(w, team=1)[2]
    b = run(w, team=1)[2]
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_team_gives_back_when_barrier_gives_up(case):
    w = case
    ip, ix, d_fail, st, steps = run(w, team=1, team_spin=1)
    _, _, d_solo, _, _ = run(w, team=0)
    print(f"[team] forced give-up: pass {st['ms_total']:.1f} ms, fp64 fallback {st['n_fallback']}, "
          f"teams gave up {st['n_team_giveup']}, handed back {st['n_team_handback']}", flush=True)
    assert st["n_unconverged"] == 0 and np.all(steps != 0)
    assert st["n_team_giveup"] > 0 and st["n_team_handback"] > 0
    assert np.abs(d_fail - d_solo).max() < TOL_T


def _local_pass(w, team, team_spin=0):
    ha.set_options(method="polar", tile=0, team=team, team_spin=team_spin)
    s = w.sls
    plan = ha.Plan(w.X.shape[0], s.nv, s.indices, s.offsets, s.dists, s.radius)
    try:
        plan.load(w.X, w.Y)
        plan.execute()
        st = plan.stats()
        steps, flags = plan.steps()
        local = plan.local(np.arange(s.count))
    finally:
        plan.close()
    return st, steps, flags, local


@pytest.mark.parametrize("team,team_spin", [(1, 0), (0, 0), (1, 1)], ids=["teams", "solo", "teams_give_up"])
def test_rank_deficient_fp32_teams_one_tile(team, team_spin):
    """Rank-deficient and nearly square searchlights (k / (n - 1) >= 0.9: the
    order-4 mesh with n = 16, k 14-19 of 13-19) take the FP32 team kernel with 64 x 64 tiles; at kp 32 one tile per
    product, so member 0 forms every product and the others read it back. Every
    local transform of a rank-deficient searchlight (k >= n) is an optimal
    orthogonal Procrustes solution (Q^T Q = I, Q^T A symmetric positive
    semidefinite, minimal residual) to 1e-4, with teams, one workgroup each, and
    teams that give up; the full-rank ones match the oracle, k = n - 1 (the most
    ill-conditioned full-rank case) and k = 13 (bf16x3) included."""
    w = synth.make_workload("c0", n=16)
    s = w.sls
    nt = w.X.shape[0]
    deficient = np.flatnonzero(s.sizes >= nt)
    assert 0 < len(deficient) < s.count
    st, steps, flags, local = _local_pass(w, team, team_spin)
    print(f"[team fp32] team {team} spin {team_spin}: {len(deficient)} rank-deficient, fp64 {st['n_fallback']}, "
          f"gave up {st['n_team_giveup']}, handed back {st['n_team_handback']}", flush=True)
    assert st["n_unconverged"] == 0 and not flags.any() and np.all(steps != 0)
    if team and team_spin:
        assert st["n_team_giveup"] > 0 and st["n_team_handback"] > 0
    elif team:
        assert st["n_team_giveup"] == 0
    for i in range(s.count):
        l = s.indices[s.offsets[i]:s.offsets[i + 1]]
        Xs, Ys = w.X[:, l], w.Y[:, l]
        T = local[i].astype(np.float64)
        if s.sizes[i] >= nt:
            k = len(l)
            assert np.abs(T.T @ T - np.eye(k)).max() < TOL_T, i
            H = T.T @ (Xs.T @ Ys)
            scale = max(1.0, np.abs(H).max())
            assert np.abs(H - H.T).max() < TOL_T * scale, i
            assert np.linalg.eigvalsh(0.5 * (H + H.T)).min() > -TOL_T * scale, i
            r, ro = np.linalg.norm(Xs @ T - Ys), np.linalg.norm(Xs @ O.procrustes(Xs, Ys) - Ys)
            assert abs(r - ro) <= TOL_T * max(1.0, ro), i
        else:
            # k = 14 and 15 (k / (n - 1) >= 0.9) run on FP32 products: the bar;
            # k = 13 (0.87) stays on bf16x3, whose error at this size and
            # conditioning (small k: entries near 0.28) measured 2.2e-4
            tol = TOL_T if 10 * k_of(s, i) >= 9 * (nt - 1) else TOL_BF16X3_SMALL
            assert np.abs(T - O.procrustes(Xs, Ys)).max() < tol, i
