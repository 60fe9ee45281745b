"""Pins the FP64 MLP oracle: its hand-derived backward pass against central
finite differences, and the layout against an explicit per-sample loop."""

import numpy as np

from oracle import mlp_oracle as M


def _data(n=40, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, M.IN)) * 2 + 1
    Y = rng.standard_normal((n, M.OUT))
    norm = np.concatenate([X.mean(0), X.std(0)])
    return X, Y, norm


def test_gradient_finite_differences():
    X, Y, norm = _data()
    p = M.init_params(1).astype(np.float64)
    p[M.NPARAM - 70:] += 0.05  # non-zero biases
    loss, g = M.loss_grad(p, norm, X, Y, w=(3.0, 0.5))
    rng = np.random.default_rng(2)
    idx = rng.choice(M.NPARAM, 60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros(M.NPARAM)
        e[i] = h
        lp, _ = M.loss_grad(p + e, norm, X, Y, w=(3.0, 0.5))
        lm, _ = M.loss_grad(p - e, norm, X, Y, w=(3.0, 0.5))
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_forward_matches_loop():
    X, Y, norm = _data(5)
    p = M.init_params(3).astype(np.float64)
    P = M.unpack(p)
    out, _ = M.forward(p, norm, X)
    for s in range(5):
        x = (X[s] - norm[:M.IN]) / norm[M.IN:]
        h1 = [M.gelu(sum(x[i] * P["W1"][i, j] for i in range(M.IN)) + P["b1"][j]) for j in range(M.HID)]
        h2 = [M.gelu(sum(h1[i] * P["W2"][i, j] for i in range(M.HID)) + P["b2"][j]) for j in range(M.HID)]
        o = [sum(h2[i] * P["W3"][i, j] for i in range(M.HID)) + P["b3"][j] for j in range(M.OUT)]
        assert np.allclose(out[s], o, rtol=1e-12, atol=1e-12)


def test_gelu_grad():
    x = np.linspace(-5, 5, 101)
    h = 1e-6
    assert np.allclose((M.gelu(x + h) - M.gelu(x - h)) / (2 * h), M.gelu_grad(x), atol=1e-8)
