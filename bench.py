"""Searchlight Procrustes throughput on one MI355X.

A step is one full pass of the hot path over the synthetic workload with X, Y
and the searchlights already resident in HBM: distance weights, cross-
covariances, local polar factors (plus FP64 fallback), and the deterministic
scatter into the CSC transform. Prints one JSON line (searchlights/s).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2] [--method auto]

Two more modes time the paths beside it, one JSON line each: --mode ridge (the
searchlight ridge pass, ha_searchlight_ridge, on the same workload) and --mode
template (the 16-subject searchlight template: order-5 icosphere,
n = 400, 20 mm, level2 = 2, through ha_searchlight_template_flat).
"""

import os

# the CPU baseline runs one searchlight per OpenMP thread, each with a
# single-threaded BLAS (oracle/cpu_ref.c binds its threads to the process's CPUs)
for _v in ("OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS", "OMP_NUM_THREADS_BLAS"):
    os.environ.setdefault(_v, "1")

import argparse
import json
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import hyperalignment_amd as ha  # noqa: E402
from hyperalignment_amd import synth  # noqa: E402

PEAK_BF16 = 2.5e15
PEAK_FP64 = 78.6e12
PEAK_HBM = 8.0e12

WORKLOADS = {
    "c0": "order-4 icosphere (2,562 vertices), 15 mm searchlights, n=200",
    "c1": "order-5 icosphere (10,242 vertices), 20 mm searchlights, n=300",
    "c2": "order-6 icosphere (40,962 vertices), 20 mm searchlights, n=500 (full hemisphere)",
    "c2tail": "order-6 icosphere (40,962 vertices), per-centre radii for sizes 37-530 with a long tail "
              "(median 90), n=500",
}


def cpu_topology():
    """The host CPU's model name, logical CPUs and physical cores (distinct
    (physical id, core id) pairs of /proc/cpuinfo)."""
    model, logical, cores = "unknown", 0, set()
    phys = core = None
    try:
        for line in open("/proc/cpuinfo"):
            key, _, val = line.partition(":")
            key, val = key.strip(), val.strip()
            if key == "model name":
                model = val
                logical += 1
            elif key == "physical id":
                phys = val
            elif key == "core id":
                core = val
            elif not key and phys is not None:
                cores.add((phys, core))
                phys = core = None
        if phys is not None:
            cores.add((phys, core))
    except OSError:
        pass
    return model, logical, (len(cores) or logical)


def cpu_baseline(w, seconds, threads, repeats=3, reflection=True, cpus=None):
    """The C/OpenMP restatement of the reference's CPU FP64 path
    (oracle/cpu_ref.c: ha_searchlight.c:950-1106 with the local solve of
    ha_searchlight.c:578-659, LAPACK dgesdd / dgesvd and BLAS dgemm, one
    searchlight per thread) on bounded samples: `repeats` runs over random
    searchlights in batches of 4 per thread, each for about seconds / repeats
    of wall time; the median rate of the runs (the reference reports medians of
    repeats, README.md:188). Returns (median rate, per-run rates, searchlights,
    transforms)."""
    from oracle import cpu_ref
    s = w.sls
    rng = np.random.default_rng(0)
    perm = rng.permutation(s.count)
    done, Ts, rates = [], [], []
    from threadpoolctl import threadpool_limits
    with threadpool_limits(limits=1, user_api="blas"):
        for _ in range(repeats):
            n0 = len(done)
            t0 = time.perf_counter()
            while time.perf_counter() - t0 < seconds / repeats and len(done) < s.count:
                batch = perm[len(done):len(done) + 4 * threads]
                Ts.extend(cpu_ref.procrustes_searchlights(w.X, w.Y, s.indices, s.offsets, batch, threads,
                                                          reflection=reflection, cpus=cpus))
                done.extend(batch.tolist())
            rates.append((len(done) - n0) / (time.perf_counter() - t0))
    return float(np.median(rates)), rates, np.array(done), Ts


def end_to_end(w):
    """The whole call a reference user makes, through the C-ABI, from host
    arrays to the host CSC transform: searchlight weights
    (ha_searchlight_weights), the CSC pattern (ha_sparse_init) and
    ha_searchlight_procrustes (plan, FP64 upload of X and Y, tile choice on a
    first pass, the pass, download of the CSC data); the reference times this
    call around hyperalignment_c.py:198-213 (benchmark_neuroboros.py:161-201).
    Two calls: the first on a fresh library state, the second after it."""
    s = w.sls
    lists, dl = s.lists(), s.dist_lists()
    out = []
    for _ in range(2):
        t0 = time.perf_counter()
        wts = ha.searchlight_weights(lists, dl, s.radius)
        t1 = time.perf_counter()
        m = ha.sparse_init(lists, s.nv)
        t2 = time.perf_counter()
        ha.searchlight_procrustes_csc(w.X, w.Y, lists, m, weights=wts)
        t3 = time.perf_counter()
        out.append({"s": t3 - t0, "weights_s": t1 - t0, "pattern_s": t2 - t1, "procrustes_s": t3 - t2})
        nnz = m.nnz
        del m, wts
    return {"first": out[0], "second": out[1], "nnz": nnz,
            "includes": "weights, CSC pattern, plan and index build, FP64 upload of X and Y, tile choice, "
                        "the pass, CSC download (host arrays in and out)"}


def comparator(w, args, every):
    """The plain paths the learned one is measured against, on the same inputs:
    the plain polar iteration (fixed scale ||A||_F and the 1e-7 schedule) over
    the whole workload, the default method without reflections, and the FP32
    one-sided Jacobi SVD path on every `every`-th searchlight, its rate
    extrapolated per searchlight."""
    from hyperalignment_amd import learn
    s = w.sls
    nt = w.X.shape[0]
    res = {}
    ha.set_options(method="polar")
    plan = ha.Plan(nt, s.nv, s.indices, s.offsets, s.dists, s.radius)
    plan.load(w.X, w.Y)
    plan.execute()
    t0 = time.perf_counter()
    plan.execute()
    dt = time.perf_counter() - t0
    st = plan.stats()
    res["plain_polar"] = {"value": s.count / dt, "hemisphere_s": dt,
                          "mean_steps": st["steps_total"] / max(1, s.count),
                          "fp64_fallback": st["n_fallback"]}
    plan.close()
    print(f"[bench] plain polar {1e3 * dt:.1f} ms", file=sys.stderr, flush=True)
    # the default method without reflections (isReflection = false): the polar
    # pass plus the det(Q) < 0 fix
    ha.set_options(method=args.method)
    plan = ha.Plan(nt, s.nv, s.indices, s.offsets, s.dists, s.radius)
    plan.load(w.X, w.Y)
    plan.execute(reflection=False)
    t0 = time.perf_counter()
    plan.execute(reflection=False)
    dt = time.perf_counter() - t0
    st = plan.stats()
    res["no_reflection"] = {"value": s.count / dt, "hemisphere_s": dt, "fp64_fallback": st["n_fallback"]}
    plan.close()
    print(f"[bench] without reflections {1e3 * dt:.1f} ms", file=sys.stderr, flush=True)
    which = np.arange(0, s.count, every)
    idx, off, d = learn.sub_searchlights(s, which)
    ha.set_options(method="jacobi32")
    plan = ha.Plan(nt, s.nv, idx, off, d, s.radius)
    plan.load(w.X, w.Y)
    t0 = time.perf_counter()
    plan.execute()
    dt = time.perf_counter() - t0
    st = plan.stats()
    res["jacobi32"] = {"value": len(which) / dt, "hemisphere_s_est": s.count * dt / len(which),
                       "sample": f"every {every}th searchlight ({len(which)}), one pass",
                       "note": "unoptimised comparator: scalar one-sided Jacobi, 16 lanes per column "
                               "pair, columns in HBM, one workgroup barrier per tournament round",
                       "mean_sweeps": st["steps_total"] / max(1, len(which))}
    plan.close()
    print(f"[bench] jacobi32 sample {1e3 * dt:.1f} ms", file=sys.stderr, flush=True)
    ha.set_options(method=args.method)
    return res


def copy_calibration():
    """One device-to-device copy of 2^28 float32 (1 GiB read, 1 GiB written),
    the known-size kernel the FETCH_SIZE / WRITE_SIZE counters are scaled by."""
    import torch
    x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": 4 << 28}))


def pmc_traffic(config, method, tiles_used, team):
    """HBM bytes per iteration pass from the committed rocprofv3 counter summary
    (FETCH_SIZE + WRITE_SIZE, calibrated against a copy kernel), when one exists
    for this workload, method, iteration tile and team setting."""
    path = os.path.join(ROOT, "profiles", f"k_polar_{config}_traffic.txt")
    if not os.path.exists(path):
        return None
    kv = {}
    for line in open(path):
        if "=" in line and not line.startswith("#"):
            k, v = line.split("=", 1)
            kv[k.strip()] = v.strip()
    if kv.get("method") != method or kv.get("tile") != tiles_used or kv.get("team") != str(team):
        return None
    return float(kv["traffic_bytes"])


def kp_bucket(kp):
    """Index of the iteration-tile bucket of a padded searchlight size (kp <= 64,
    128, 256, 384, 512, larger), as the library's autotuner groups them."""
    for b, edge in enumerate((64, 128, 256, 384, 512)):
        if kp <= edge:
            return b
    return 5


def one_blas_thread():
    """numpy's BLAS limited to one thread (whatever the environment sets): the
    single-core CPU samples below are many small factorizations, which a
    multithreaded BLAS slows down."""
    from threadpoolctl import threadpool_limits
    return threadpool_limits(limits=1, user_api="blas")


def bench_ridge(args):
    """The searchlight ridge pass (ha_searchlight_ridge, ha_searchlight.c:1378-1429): per
    searchlight betas = (X_sl^T X_sl + alpha I)^-1 X_sl^T Y_sl in FP64, scattered like the
    Procrustes transforms. Roofline: the solve kernel's 3 n k^2 + 7 k^3 / 3 FP64 flops per
    searchlight against the FP64 MFMA peak."""
    w = synth.make_workload(args.config)
    s = w.sls
    nt = w.X.shape[0]
    ha.init()
    plan = ha.Plan(nt, s.nv, s.indices, s.offsets, s.dists, s.radius)
    plan.load(w.X, w.Y)
    for i in range(args.warmup):
        plan.ridge(args.alpha)
        print(f"[bench] ridge warmup {i} {plan.stats()['ms_total']:.1f} ms", file=sys.stderr, flush=True)
    solve_ms = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        plan.ridge(args.alpha)
        st = plan.stats()
        solve_ms.append(st["ms_polar"])
        print(f"[bench] ridge step {i} {st['ms_total']:.1f} ms (solve {st['ms_polar']:.1f} ms)", file=sys.stderr,
              flush=True)
    dt = time.perf_counter() - t0
    k = s.sizes.astype(np.float64)
    flops = float((3.0 * nt * k * k + 7.0 * k ** 3 / 3.0).sum())
    kms = float(np.mean(solve_ms))
    achieved = flops / (kms * 1e-3)
    roofline = {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP64, "traffic": None, "kernel": "k_ridge64", "kernel_ms": kms,
                "flops_per_launch": flops,
                "kernel_timing": "HIP events on the library stream around the solve launch of every timed pass"}
    cpu = parity = None
    if not args.no_cpu:
        from oracle import ha_oracle as O
        rng = np.random.default_rng(11)
        order = rng.permutation(s.count)
        lists = s.lists()
        done, betas = [], []
        with one_blas_thread():
            t0 = time.perf_counter()
            while time.perf_counter() - t0 < args.cpu_seconds / 2 and len(done) < s.count:
                c = int(order[len(done)])
                betas.append(O.ridge(w.X[:, lists[c]], w.Y[:, lists[c]], args.alpha))
                done.append(c)
            rate = len(done) / (time.perf_counter() - t0)
        cpu = {"value": rate, "unit": "searchlights/s", "cores": 1, "kind": "port",
               "sample": f"{len(done)} random searchlights of {args.config}: the oracle's FP64 ridge through the "
                         f"SVD of X_sl (ha_ridge.c:13-54, ridge.py:7-12), numpy with a single-threaded BLAS; "
                         f"hemisphere estimate {s.count / rate:.0f} s"}
        loc = plan.local(np.array(done, dtype=np.int32))
        errs = [float(np.abs(a - b).max() / max(1e-30, np.abs(b).max())) for a, b in zip(loc, betas)]
        parity = {"local_max_rel": max(errs), "local_median_rel": float(np.median(errs)), "searchlights": len(errs),
                  "note": "max |GPU betas (stored FP32) - oracle| over the largest |beta| of each searchlight"}
    out = {"metric": "searchlights/s (full-hemisphere searchlight ridge pass)", "value": s.count * args.steps / dt,
           "unit": "searchlights/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": 1e3 * dt / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"{args.config}: {WORKLOADS[args.config]}", "alpha": args.alpha,
                      "stage_ms": {"ms_weights": st["ms_weights"], "ms_solve": st["ms_polar"],
                                   "ms_scatter": st["ms_scatter"]},
                      "parity_vs_fp64": parity},
           "roofline": roofline, "cpu_baseline": cpu}
    print(json.dumps(out))
    plan.close()


def bench_template(args):
    """The 16-subject searchlight template (ha_searchlight_template_flat; ha_template.c:14-114
    per searchlight, combined as searchlight.py:215-236): 16 subjects, each call uploads them,
    runs the template kernel and the FP64 gather, and downloads the template."""
    dms, sls = synth.template_subjects(n_subj=16, n=400, order=5, radius=20.0, seed=0)
    ha.init()

    def run():
        return ha.searchlight_template_flat(dms, sls.indices, sls.offsets, sls.dists, sls.radius, level2_iter=2)

    for i in range(args.warmup):
        t0 = time.perf_counter()
        run()
        print(f"[bench] template warmup {i} {time.perf_counter() - t0:.2f} s", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    kms, flops = [], []
    for i in range(args.steps):
        tpl = run()
        st = ha.last_template_stats()
        kms.append(st.ms_polar)
        flops.append(st.gemm_flops)
    dt = time.perf_counter() - t0
    kms_avg, flops_avg = float(np.mean(kms)), float(np.mean(flops))
    achieved = flops_avg / (kms_avg * 1e-3)
    roofline = {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP64, "traffic": None, "kernel": "k_template<false>", "kernel_ms": kms_avg,
                "flops_per_launch": flops_avg,
                "flops": "FP64 algorithmic: 2 n k^2 per solve (A = D_s^T R) and 4 k^3 per polar step "
                         "(X^T X and G^2 symmetric, X P), from each searchlight's step count; the FP32 "
                         "products D_s Q are not counted",
                "kernel_timing": "HIP events on the library stream around the template launch of every timed call"}
    cpu = None
    if not args.no_cpu:
        from oracle import ha_oracle as O
        lists = sls.lists()
        rng = np.random.default_rng(12)
        order = rng.permutation(sls.count)
        n = 0
        with one_blas_thread():
            t0 = time.perf_counter()
            while time.perf_counter() - t0 < args.cpu_seconds / 2 and n < sls.count:
                c = lists[int(order[n])]
                O.template_procrustes([d[:, c] for d in dms], level2_iter=2)
                n += 1
            rate = n / (time.perf_counter() - t0)
        cpu = {"value": rate, "unit": "searchlights/s", "cores": 1, "kind": "port",
               "sample": f"{n} random searchlights' local templates (the oracle's FP64 chain with LAPACK SVDs, "
                         f"ha_template.c:14-114), numpy with a single-threaded BLAS; whole template estimate "
                         f"{sls.count / rate:.0f} s"}
    out = {"metric": "searchlights/s (16-subject searchlight template, 47 Procrustes solves per searchlight)",
           "value": sls.count * args.steps / dt, "unit": "searchlights/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": 1e3 * dt / args.steps, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64 (A and polar factors) + f32 (storage, D Q)",
           "data": "synthetic",
           "config": {"workload": "template: 16 subjects, order-5 icosphere (10,242 vertices), n = 400, 20 mm "
                                  "searchlights, level2 = 2, z-scored",
                      "timed": "each step a whole ha_searchlight_template_flat call (host data in, template out)",
                      "finite": bool(np.all(np.isfinite(tpl))),
                      "parity_vs_fp64": "tests/test_gpu_template.py (16 columns within 1.2e-7 of the largest "
                                        "magnitude)"},
           "roofline": roofline,
           "cpu_baseline": cpu}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c2", choices=sorted(WORKLOADS))
    ap.add_argument("--method", default="auto", choices=sorted(ha.METHODS))
    ap.add_argument("--cpu-seconds", type=float, default=18.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--no-compare", action="store_true", help="skip the plain polar / Jacobi32 comparison")
    ap.add_argument("--no-e2e", action="store_true", help="skip the end-to-end C-ABI call timing")
    ap.add_argument("--eager", action="store_true",
                    help="time eager passes instead of replays of the captured hipGraph (the default: the "
                         "library captures a plan's pass on its second run and replays it)")
    ap.add_argument("--compare-every", type=int, default=40)
    ap.add_argument("--no-reflection", action="store_true",
                    help="time passes with isReflection = false (the det(Q) < 0 fix, ha_searchlight.c:612-632)")
    ap.add_argument("--mode", default="procrustes", choices=["procrustes", "ridge", "template"],
                    help="the searchlight Procrustes pass (the headline), the ridge pass, or the 16-subject template")
    ap.add_argument("--alpha", type=float, default=1.0, help="ridge penalty (--mode ridge)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device-to-device copy (reference for the HBM byte counters)")
    args = ap.parse_args()
    refl = not args.no_reflection
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              f"(the searchlight pass shards by owner-computes column blocks, not run here)", file=sys.stderr)
        sys.exit(2)
    if args.mode == "ridge":
        bench_ridge(args)
        return
    if args.mode == "template":
        bench_template(args)
        return

    t0 = time.perf_counter()
    w = synth.make_workload(args.config)
    s = w.sls
    nt = w.X.shape[0]
    print(f"[bench] workload {args.config}: {s.count} searchlights, k {s.sizes.min()}-{s.sizes.max()}, "
          f"n={nt} (generated in {time.perf_counter() - t0:.1f}s)", file=sys.stderr, flush=True)
    ha.init()
    ha.set_options(method=args.method, tile=args.tile, use_graph=0 if args.eager else 1)
    from hyperalignment_amd import learn
    if os.path.exists(learn.MODEL_PATH):
        learn.install(learn.MODEL_PATH)
    plan = ha.Plan(nt, s.nv, s.indices, s.offsets, s.dists, s.radius)
    plan.load(w.X, w.Y)
    print(f"[bench] plan ready: nnz {plan.nnz()}, pattern {plan.stats()['ms_pattern']:.1f} ms",
          file=sys.stderr, flush=True)
    for i in range(args.warmup):
        plan.execute(reflection=refl)
        print(f"[bench] warmup {i} {plan.stats()['ms_total']:.1f} ms"
              f"{' (graph replay)' if plan.stats()['graph_replay'] else ''}", file=sys.stderr, flush=True)
    if not args.eager and not plan.stats()["graph_replay"]:
        # the timed passes replay a graph captured outside the timed region
        plan.execute(reflection=refl)
        print("[bench] capture pass", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    polar_ms, replays = [], 0
    for i in range(args.steps):
        plan.execute(reflection=refl)
        st = plan.stats()
        polar_ms.append(st["kernel_ms_polar_avg"])
        replays += st["graph_replay"]
        print(f"[bench] step {i} {st['ms_total']:.1f} ms (iteration {st['kernel_ms_polar_avg']:.1f} ms)",
              file=sys.stderr, flush=True)
    dt = time.perf_counter() - t0
    st = plan.stats()
    ms_step = 1e3 * dt / args.steps
    value = s.count * args.steps / dt

    # dominant kernel: the local-transform kernel (cross-covariance + polar iteration)
    polar_avg = float(np.mean(polar_ms))
    kernel_timing = ("HIP events recorded inside the captured hipGraph on the library stream, around the "
                     "iteration launches, in every timed replay" if replays else
                     "HIP events on the library stream around the iteration, inside the timed passes")
    if not min(polar_ms) > 0:
        # a captured pass without its stage events: one eager pass after the timed region
        ha.set_options(use_graph=0)
        plan.execute(reflection=refl)
        polar_avg = plan.stats()["kernel_ms_polar_avg"]
        ha.set_options(use_graph=0 if args.eager else 1)
        kernel_timing = "HIP events around the iteration of one eager pass after the timed region"
    achieved = st["gemm_flops"] / (polar_avg * 1e-3)
    method = ha.resolved_method(args.method, int(s.sizes.max()))
    tiles = None
    tiles_used = None
    if method in ("polar", "polar_learned"):
        edges = ["kp<=64", "kp<=128", "kp<=256", "kp<=384", "kp<=512", "kp>512"]
        bt = [int(t) for t in plan.tiles()]
        tiles = dict(zip(edges, bt))
        kps = set(((s.sizes + 31) // 32 * 32).tolist())
        tiles_used = "+".join(str(t) for t in sorted({bt[kp_bucket(kp)] for kp in kps}))
    team = 0
    if tiles_used:
        tn = plan.tuning()
        team = int(any(tn[kp_bucket(kp)]["teams"] for kp in set(((s.sizes + 31) // 32 * 32).tolist())))
    kernel = "k_polar" + (f"<{tiles_used}>" if tiles_used else "")
    if team:
        kernel = "k_polar_grp (teams of 4, 128 x 128 tiles) + " + kernel
    roofline = {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK_BF16 / 1e12, "unit": "TFLOP/s",
                "frac": achieved / PEAK_BF16, "traffic": pmc_traffic(args.config, method, tiles_used, team),
                "kernel": kernel,
                "kernel_ms": polar_avg, "flops_per_launch": st["gemm_flops"], "kernel_timing": kernel_timing}
    stats_main = st
    # apply and score the transform on the training data (Yhat = X T, vertex-wise r)
    t0 = time.perf_counter()
    _, r = plan.evaluate(w.X, w.Y, want_yhat=False)
    ev_st = plan.stats()
    nnz = st["nnz"]
    ev_bytes = 2 * nt * s.nv * 8 + nnz * 12  # X read once, T (data + index), Yhat written
    evaluate = {"ms_apply": ev_st["ms_apply"], "ms_corr": ev_st["ms_corr"],
                "apply_algorithmic_GBs": ev_bytes / (ev_st["ms_apply"] * 1e-3) / 1e9,
                "r_mean": float(np.nanmean(r)), "host_s": time.perf_counter() - t0}
    tuning = plan.tuning()
    other = None
    if True:
        # the same pass launched the other way (eager when the timed passes were graph replays)
        ha.set_options(use_graph=1 if args.eager else 0)
        plan.execute(reflection=refl)
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            plan.execute(reflection=refl)
            ts.append(time.perf_counter() - t0)
        ha.set_options(use_graph=0 if args.eager else 1)
        other = {"launch": "hipGraph replay" if args.eager else "eager",
                 "s_per_pass": float(np.median(ts)), "runs": ts, "gpu_ms_last": plan.stats()["ms_total"]}
        print(f"[bench] {other['launch']} {1e3 * other['s_per_pass']:.1f} ms", file=sys.stderr, flush=True)
    e2e = None
    if not args.no_e2e:
        e2e = end_to_end(w)
        print(f"[bench] end to end {e2e['first']['s']:.2f} s, then {e2e['second']['s']:.2f} s", file=sys.stderr,
              flush=True)
    compare = None
    if not args.no_compare and method in ("polar", "polar_learned"):
        compare = comparator(w, args, args.compare_every)
        compare["speedup_vs_jacobi32"] = value / compare["jacobi32"]["value"]
        compare["speedup_vs_plain_polar"] = value / compare["plain_polar"]["value"]

    # parity on a sample of searchlights against the FP64 oracle
    st = stats_main
    cpu = None
    parity = None
    if not args.no_cpu:
        from oracle import cpu_ref
        avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        model, ncpu, ncores = cpu_topology()
        # one thread per physical core this process may use, each bound to it
        cpus = cpu_ref.core_cpus()
        threads = len(cpus)
        with one_blas_thread():  # one searchlight per OpenMP thread, each with a single-threaded BLAS
            rate, rates, which, Ts = cpu_baseline(w, args.cpu_seconds, threads, reflection=refl, cpus=cpus)
        cpu = {"value": rate, "unit": "searchlights/s", "cores": threads, "kind": "port",
               "sample": f"median of 3 runs ({', '.join(f'{r:.1f}' for r in rates)} searchlights/s) over "
                         f"{len(which)} random searchlights of {args.config} (about {args.cpu_seconds:.0f} s in "
                         f"all): the C/OpenMP restatement of the reference's CPU FP64 path (oracle/cpu_ref.c: "
                         f"gather, dgemm, dgesdd, U Vt), one searchlight per OpenMP thread with a single-threaded "
                         f"BLAS, {threads} threads each bound to one CPU of a distinct core ({avail} CPUs available to the "
                         f"process; "
                         f"{ncpu} logical CPUs, {ncores} physical cores, {model}); hemisphere estimate "
                         f"{s.count / rate:.1f} s",
               "cpu_model": model, "logical_cpus": ncpu, "physical_cores": ncores, "cpus_available": avail,
               "runs": rates}
        loc = plan.local(which)
        errs = [float(np.abs(a - b).max()) for a, b in zip(loc, Ts)]
        fro = [float(np.linalg.norm(a - b) / np.linalg.norm(b)) for a, b in zip(loc, Ts)]
        parity = {"local_maxabs": max(errs), "local_median": float(np.median(errs)),
                  "local_rel_frobenius_max": max(fro), "local_rel_frobenius_median": float(np.median(fro)),
                  "searchlights": len(errs)}

    out = {
        "metric": "searchlights/s (full-hemisphere searchlight Procrustes pass)",
        "value": value,
        "unit": "searchlights/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3" if method in ("polar", "polar_learned") else method,
        "data": "synthetic",
        "config": {"workload": f"{args.config}: {WORKLOADS[args.config]}", "method": method,
                   "isReflection": refl,
                   "launch": (f"hipGraph replay (captured in an untimed pass; {replays} of {args.steps} timed passes "
                              f"were replays)" if not args.eager else "eager"),
                   "launch_compare": other,
                   "searchlights": s.count, "hemisphere_s": dt / args.steps,
                   "pattern_ms": st["ms_pattern"], "nnz": st["nnz"], "fp64_fallback": st["n_fallback"],
                   "unconverged": st["n_unconverged"],
                   "mean_steps": st["steps_total"] / max(1, s.count),
                   "steps_hist": {str(i): int(v) for i, v in enumerate(st["steps_hist"]) if v},
                   "orthogonality_check": {"checked": st["n_orth_checked"], "sent_to_fp64": st["n_orth_flagged"],
                                           "max_fro": st["orth_max"], "ms": st["ms_orth"]},
                   "teams": {"gave_up": st["n_team_giveup"], "searchlights_handed_back": st["n_team_handback"]},
                   "stage_ms": {k: st[k] for k in ("ms_weights", "ms_crosscov", "ms_polar", "ms_fallback",
                                                   "ms_scatter")},
                   "iteration_tile": tiles, "autotune": tuning, "parity_vs_fp64": parity, "compare": compare,
                   "end_to_end": e2e,
                   "evaluate": evaluate},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    plan.close()


if __name__ == "__main__":
    main()
