"""Learned spectral scaling of the polar iteration: data, training, evaluation.

The polar kernel computes 12 cheap features of each cross-covariance A in its
prologue (size, ||A||_F, trace and diagonal statistics, three power-iteration
estimates of sigma_max, n_samples) and feeds them to a 12 -> 64 -> 64 -> 2 GELU
MLP that predicts

    y0 = log(sigma_max / ||A||_F)      (the scale of X0 = A / scale)
    y1 = log10(sigma_min / sigma_max)  (the schedule bucket, i.e. iteration count)

Labels are the exact spectra of the FP32 one-sided Jacobi SVD path. Training
uses the library's own GPU forward/backward/Adam kernels. The kernel scales A by
the predicted sigma_max raised by a margin and picks the schedule (the number
of steps) from the predicted sigma_min / sigma_max lowered by a margin, both
calibrated on held-out rows. A wrong prediction costs steps, never accuracy: an
ell too small costs about half a step per decade, one too large a restart of the
schedule, and a scale below sigma_max makes the attempt diverge, after which the
kernel redoes the searchlight as the plain iteration.

Workload grid: the full-hemisphere mesh (order-6 icosphere, 20 mm
searchlights) with seeds 0-7, n_samples in {200, 300, 500} and noise sigma in
{0.1, 0.3, 1.0}; evaluation on held-out seeds.

  python -m hyperalignment_amd.learn train --out polar_mlp.npz
  python -m hyperalignment_amd.learn eval --model hyperalignment_amd/model/polar_mlp.npz
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import synth
import hyperalignment_amd as ha

MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "model", "polar_mlp.npz")
TRAIN_SEEDS = list(range(8))
EVAL_SEEDS = [100, 101]
NS = [200, 300, 500]
NOISES = [0.1, 0.3, 1.0]


def log(msg):
    print(f"[learn] {msg}", file=sys.stderr, flush=True)


def init_params(seed=0):
    """Glorot-uniform weights, zero biases, in the parameter layout of the kernels."""
    rng = np.random.default_rng(seed)
    parts = []
    for shape in [(12, 64), (64,), (64, 64), (64,), (64, 2), (2,)]:
        if len(shape) == 2:
            lim = np.sqrt(6.0 / (shape[0] + shape[1]))
            parts.append(rng.uniform(-lim, lim, size=shape).ravel())
        else:
            parts.append(np.zeros(shape))
    return np.concatenate(parts).astype(np.float32)


def sub_searchlights(sls: synth.Searchlights, which):
    """Flat arrays of a subset of searchlights (same vertex ids)."""
    lists = sls.lists()
    dl = sls.dist_lists()
    sizes = np.array([len(lists[s]) for s in which], dtype=np.int64)
    off = np.zeros(len(which) + 1, dtype=np.int64)
    np.cumsum(sizes, out=off[1:])
    idx = np.concatenate([lists[s] for s in which]).astype(np.int32)
    d = np.concatenate([dl[s] for s in which]).astype(np.float64)
    return idx, off.astype(np.int32), d


def dataset_rows(w: synth.Workload, sample: int, rng):
    """Features and labels of `sample` random searchlights of one workload."""
    which = np.sort(rng.choice(w.sls.count, min(sample, w.sls.count), replace=False))
    idx, off, d = sub_searchlights(w.sls, which)
    plan = ha.Plan(w.X.shape[0], w.sls.nv, idx, off, d, w.sls.radius)
    try:
        plan.load(w.X, w.Y)
        f, sig = plan.spectra(np.arange(len(which)))
    finally:
        plan.close()
    fro = np.exp(f[:, 1].astype(np.float64))
    y0 = np.log(sig[:, 0].astype(np.float64) / fro)
    y1 = np.log10(np.maximum(sig[:, 1].astype(np.float64), 1e-30) / sig[:, 0])
    return f, np.stack([y0, y1], axis=1).astype(np.float32), which


ELL_FALLBACK_LOG10 = -6.0  # training rows: sigma_min / sigma_max >= 1e-6


def usable(f, y):
    """Rows the polar iteration handles: full rank (k < n_samples) and
    sigma_min / sigma_max >= 1e-6; the rest go to the FP64 Jacobi fallback."""
    k = np.exp(f[:, 0].astype(np.float64))
    nt = np.exp(f[:, 10].astype(np.float64))
    return (k < nt - 0.5) & (y[:, 1] > ELL_FALLBACK_LOG10)


def collect(seeds, ns, noises, sample, rng_seed=0):
    rng = np.random.default_rng(rng_seed)
    X, Y, meta = [], [], []
    for seed in seeds:
        for n in ns:
            for noise in noises:
                t0 = time.perf_counter()
                w = synth.make_workload("c2", seed=seed, n=n, noise=noise)
                if w.sls.sizes.min() >= n:
                    log(f"seed {seed} n {n} noise {noise}: every searchlight has k >= n "
                        f"(rank deficient, FP64 fallback); no polar rows")
                    meta.append((seed, n, noise, 0))
                    continue
                f, y, _ = dataset_rows(w, sample, rng)
                ok = usable(f, y)
                X.append(f[ok])
                Y.append(y[ok])
                meta.append((seed, n, noise, int(ok.sum())))
                log(f"seed {seed} n {n} noise {noise}: {ok.sum()}/{len(f)} rows, y1 median "
                    f"{np.median(y[ok, 1]):.2f} min {y[ok, 1].min():.2f} ({time.perf_counter() - t0:.1f}s)")
    return np.concatenate(X), np.concatenate(Y), meta


def fit(X, Y, epochs=60, batch=4096, lr=2e-3, w=(10.0, 1.0), seed=0):
    mean = X.mean(axis=0)
    std = X.std(axis=0)
    std[std < 1e-6] = 1.0
    norm = np.concatenate([mean, std]).astype(np.float32)
    params = init_params(seed)
    params, hist = ha.mlp_train(params, norm, X, Y, epochs=epochs, batch=batch, lr=lr, w=w, seed=seed)
    return params, norm, hist


def calibrate(params, norm, X, Y, q_ell=0.95):
    """Margins from held-out rows. Scale: the largest under-prediction of
    ln(sigma_max/||A||_F) plus 0.05, since an attempt whose scale is below
    sigma_max diverges and is redone. Ell: the q_ell quantile of the
    over-prediction of log10(sigma_min/sigma_max) plus 0.1 decades; the
    remaining rows restart their schedule, which costs a few steps. (A
    simulation of the schedules on C2 spectra with the validation errors puts
    the fewest mean steps near these margins.)"""
    P = ha.mlp_forward(params, norm, X).astype(np.float64)
    m0 = float(np.max(Y[:, 0] - P[:, 0])) + 0.05
    m1 = float(np.quantile(P[:, 1] - Y[:, 1], q_ell)) + 0.1
    rmse = np.sqrt(((P - Y) ** 2).mean(axis=0))
    return max(m0, 0.0), max(m1, 0.0), rmse


def save(path, params, norm, m0, m1, info):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.savez(path, params=params, norm=norm, margins=np.array([m0, m1], dtype=np.float32),
             info=np.array(json.dumps(info)))


def load(path=MODEL_PATH):
    z = np.load(path)
    return z["params"], z["norm"], [float(v) for v in z["margins"]], json.loads(str(z["info"]))


def install(path=MODEL_PATH):
    """Loads the model into the library and sets its calibrated margins."""
    params, norm, (m0, m1), info = load(path)
    ha.set_model(params, norm)
    ha.set_options(learned_margin_scale=m0, learned_margin_ell=m1)
    return info


def evaluate_dataset(w, sample, rng, model_path=MODEL_PATH):
    """Plain vs learned polar iteration on `sample` searchlights, judged by FP64 Jacobi."""
    which = np.sort(rng.choice(w.sls.count, min(sample, w.sls.count), replace=False))
    idx, off, d = sub_searchlights(w.sls, which)
    plan = ha.Plan(w.X.shape[0], w.sls.nv, idx, off, d, w.sls.radius)
    res = {}
    try:
        plan.load(w.X, w.Y)
        ref = None
        for method in ["jacobi64", "polar", "polar_learned"]:
            ha.set_options(method=method)
            plan.execute()
            st, fl = plan.steps()
            loc = plan.local(np.arange(len(which)))
            if method == "jacobi64":
                ref = loc
                continue
            err = max(float(np.abs(a - b).max()) for a, b in zip(loc, ref))
            good = st > 0
            res[method] = {"mean_steps": float(st[good].mean()) if good.any() else None,
                           "max_steps": int(st.max()), "fallback": int(fl.sum()),
                           "unconverged": int((st < 0).sum()), "maxabs_vs_fp64": err,
                           "ms": plan.stats()["ms_polar"]}
    finally:
        ha.set_options(method="auto")
        plan.close()
    return res


def cmd_train(args):
    ha.init()
    t0 = time.perf_counter()
    X, Y, meta = collect(TRAIN_SEEDS[:args.seeds], NS, NOISES, args.sample)
    log(f"collected {len(X)} rows in {time.perf_counter() - t0:.0f}s")
    rng = np.random.default_rng(1)
    perm = rng.permutation(len(X))
    nval = len(X) // 10
    va, tr = perm[:nval], perm[nval:]
    params, norm, hist = fit(X[tr], Y[tr], epochs=args.epochs)
    log(f"loss per epoch: {np.array2string(hist, precision=4)}")
    m0, m1, rmse = calibrate(params, norm, X[va], Y[va])
    log(f"validation rmse y0 {rmse[0]:.4f} y1 {rmse[1]:.4f}; margins scale {m0:.4f} ell {m1:.3f} decades")
    info = {"rows": int(len(X)), "train_rows": int(len(tr)), "val_rows": int(nval), "epochs": args.epochs,
            "loss_first": float(hist[0]), "loss_last": float(hist[-1]), "val_rmse": [float(r) for r in rmse],
            "margins": [m0, m1], "grid": {"seeds": TRAIN_SEEDS[:args.seeds], "n": NS, "noise": NOISES},
            "sample_per_dataset": args.sample}
    save(args.out, params, norm, m0, m1, info)
    np.savez(args.out.replace(".npz", "_data.npz"), X=X, Y=Y)
    log(f"saved {args.out}")


def cmd_eval(args):
    ha.init()
    info = install(args.model)
    log(f"model: {info}")
    rng = np.random.default_rng(7)
    out = []
    for seed in EVAL_SEEDS[:args.seeds]:
        for n in [int(v) for v in args.ns.split(",")]:
            for noise in NOISES:
                w = synth.make_workload("c2", seed=seed, n=n, noise=noise)
                r = evaluate_dataset(w, args.sample, rng, args.model)
                r.update({"seed": seed, "n": n, "noise": noise})
                log(json.dumps(r))
                out.append(r)
    print(json.dumps(out))


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--out", default="polar_mlp.npz")
    t.add_argument("--seeds", type=int, default=8)
    # the committed model (model/polar_mlp.npz, its info record) was trained with these
    t.add_argument("--sample", type=int, default=3000)
    t.add_argument("--epochs", type=int, default=60)
    e = sub.add_parser("eval")
    e.add_argument("--model", default=MODEL_PATH)
    e.add_argument("--seeds", type=int, default=2)
    e.add_argument("--sample", type=int, default=256)
    e.add_argument("--ns", default="200,300,500")
    args = ap.parse_args(argv)
    {"train": cmd_train, "eval": cmd_eval}[args.cmd](args)


if __name__ == "__main__":
    main()
