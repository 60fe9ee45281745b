"""GPU kernels of the learned polar scaling against the FP64 MLP oracle."""

import numpy as np
import pytest

import hyperalignment_amd as ha
from hyperalignment_amd import learn, synth
from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def hip():
    ha.init()
    yield
    ha.set_model(None, None)
    ha.set_options(method="auto")


def _data(n=1000, seed=0):
    rng = np.random.default_rng(seed)
    X = (rng.standard_normal((n, M.IN)) * 2 + 1).astype(np.float32)
    Y = np.stack([np.tanh(X[:, 0] - X[:, 3]), X[:, 1] * 0.3 - 1], axis=1).astype(np.float32)
    # the normalisation describes the generating distribution (mean 1, std 2), so it
    # stays finite for a single-row batch
    norm = np.concatenate([np.ones(M.IN), 2 * np.ones(M.IN)]).astype(np.float32)
    return X, Y, norm


def test_forward_matches_oracle():
    X, _, norm = _data()
    p = learn.init_params(5)
    got = ha.mlp_forward(p, norm, X)
    ref, _ = M.forward(p.astype(np.float64), norm.astype(np.float64), X.astype(np.float64))
    assert np.abs(got - ref).max() < 1e-4


@pytest.mark.parametrize("n", [1, 64, 1000])
def test_gradient_matches_oracle(n):
    X, Y, norm = _data(n)
    p = learn.init_params(6)
    g, loss = ha.mlp_grad(p, norm, X, Y, w=(3.0, 0.5))
    lref, gref = M.loss_grad(p.astype(np.float64), norm.astype(np.float64), X.astype(np.float64),
                             Y.astype(np.float64), w=(3.0, 0.5))
    assert abs(loss - lref) <= 1e-4 * max(1.0, abs(lref))
    scale = np.abs(gref).max()
    assert np.abs(g - gref).max() <= 2e-4 * scale


def test_gradient_deterministic():
    X, Y, norm = _data(3000)
    p = learn.init_params(7)
    a, _ = ha.mlp_grad(p, norm, X, Y)
    b, _ = ha.mlp_grad(p, norm, X, Y)
    assert np.array_equal(a, b)


def test_training_reduces_loss():
    X, Y, norm = _data(4000)
    p0 = learn.init_params(8)
    p, hist = ha.mlp_train(p0, norm, X, Y, epochs=15, batch=512, lr=3e-3)
    _, l0 = ha.mlp_grad(p0, norm, X, Y)
    _, l1 = ha.mlp_grad(p, norm, X, Y)
    assert l1 < 0.2 * l0, (l0, l1, hist)


def test_fused_inference_matches_batch_forward():
    """The polar kernel's prologue features + MLP equal the batch path."""
    w = synth.make_workload("c1")
    which = np.arange(0, w.sls.count, 97)[:64]
    idx, off, d = learn.sub_searchlights(w.sls, which)
    plan = ha.Plan(w.X.shape[0], w.sls.nv, idx, off, d, w.sls.radius)
    plan.load(w.X, w.Y)
    f, sig = plan.spectra(np.arange(len(which)))
    X, Y, norm = _data()
    norm = np.concatenate([f.mean(0), f.std(0) + 1e-3]).astype(np.float32)
    p = learn.init_params(9)
    ha.set_model(p, norm)
    ha.set_options(method="polar_learned")
    plan.execute()
    pred = plan.predictions()
    ref = ha.mlp_forward(p, norm, f)
    assert np.abs(pred - ref).max() < 1e-4
    # spectra labels: sigma_max >= power estimates, sigma_min > 0
    assert np.all(sig[:, 0] >= sig[:, 1]) and np.all(sig[:, 1] > 0)
    assert np.all(np.log(sig[:, 0]) >= f[:, 9] + f[:, 1] - 1e-3)
    plan.close()


def test_learned_scaling_gain():
    """Ablation on held-out data (seed 100, not in the training grid): the learned
    spectral scaling needs fewer polar steps than the plain schedule, with no FP64
    fallback and both within 1e-4 of the FP64 Jacobi judge."""
    learn.install()
    w = synth.make_workload("c2", seed=100, n=500, noise=0.3)
    r = learn.evaluate_dataset(w, 256, np.random.default_rng(3))
    plain, lrn = r["polar"], r["polar_learned"]
    assert lrn["fallback"] == 0 and lrn["unconverged"] == 0
    assert plain["maxabs_vs_fp64"] < 1e-4 and lrn["maxabs_vs_fp64"] < 1e-4
    assert lrn["mean_steps"] < plain["mean_steps"] - 0.5, (plain, lrn)
