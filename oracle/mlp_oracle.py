"""FP64 numpy oracle of the spectral-scaling MLP -- TEST INFRASTRUCTURE ONLY.

12 -> 64 -> 64 -> 2 with exact GELU, parameters laid out as
W1[12][64], b1[64], W2[64][64], b2[64], W3[64][2], b3[2] (input unit = row),
inputs normalised as (f - mean) / std. Loss: mean over samples of
sum_j w_j (out_j - y_j)^2. The hand-derived backward pass here is pinned by
central finite differences (tests/test_mlp_oracle.py) before it is used to
check the GPU gradient kernels. (The MLP has no counterpart in the reference.)
"""

import numpy as np
from scipy.special import erf

IN, HID, OUT = 12, 64, 2
SIZES = [("W1", (IN, HID)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)), ("W3", (HID, OUT)), ("b3", (OUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SIZES])


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def forward(p, norm, X):
    P = unpack(p)
    x = (np.asarray(X, np.float64) - norm[:IN]) / norm[IN:]
    z1 = x @ P["W1"] + P["b1"]
    h1 = gelu(z1)
    z2 = h1 @ P["W2"] + P["b2"]
    h2 = gelu(z2)
    return h2 @ P["W3"] + P["b3"], (x, z1, h1, z2, h2)


def loss_grad(p, norm, X, Y, w=(1.0, 1.0)):
    P = unpack(p)
    out, (x, z1, h1, z2, h2) = forward(p, norm, X)
    n = x.shape[0]
    wv = np.asarray(w, dtype=np.float64)
    e = out - np.asarray(Y, np.float64)
    loss = float((wv * e * e).sum() / n)
    dout = 2.0 * wv * e / n
    g = {"W3": h2.T @ dout, "b3": dout.sum(0)}
    d2 = (dout @ P["W3"].T) * gelu_grad(z2)
    g["W2"] = h1.T @ d2
    g["b2"] = d2.sum(0)
    d1 = (d2 @ P["W2"].T) * gelu_grad(z1)
    g["W1"] = x.T @ d1
    g["b1"] = d1.sum(0)
    return loss, pack(g)


def init_params(seed=0):
    """Glorot-uniform weights, zero biases."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in SIZES:
        if name.startswith("W"):
            lim = np.sqrt(6.0 / (shape[0] + shape[1]))
            d[name] = rng.uniform(-lim, lim, size=shape)
        else:
            d[name] = np.zeros(shape)
    return pack(d).astype(np.float32)
