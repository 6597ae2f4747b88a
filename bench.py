"""Benchmark: runs/s to |dlogZ| < 0.1 nat and likelihood evals/s on one MI355X.

Workload (default): D=32 multivariate-normal likelihood, prior N(0, I), seeded
mu ~ N(0, 0.5^2 I), Sigma = A A^T / D + 0.05 I; num_live_points=2048, s=4
(128 slices per chain, 1024 chains per shell); FP32 likelihood (MFMA), FP64
evidence.  One step = one batch of R independent runs (different seeds) taken
from create_init_state to log Z, in the chosen form: all R runs advanced by the
same kernel launches, shells replayed from a captured hipGraph.  The plain form
(one run at a time, host sync after every shell) is timed beside it.

Prints one JSON line (see the "roofline" and "cpu_baseline" objects for the
kernel-level measurement and the host-CPU reference).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

PEAK_FP32_MFMA = 157.3e12
PEAK_HBM = 8.0e12


def mvn_model(D, seed):
    import jaxns_amd as J
    from jaxns_amd import distributions as tfpd
    from scipy.stats import multivariate_normal
    rng = np.random.default_rng(seed)
    mu = rng.normal(0.0, 0.5, D)
    A = rng.normal(size=(D, D))
    Sig = A @ A.T / D + 0.05 * np.eye(D)
    log_Z = multivariate_normal(np.zeros(D), np.eye(D) + Sig).logpdf(mu)

    def prior_model():
        x = yield J.Prior(tfpd.Normal(loc=np.zeros(D), scale=np.ones(D)), name="x")
        return x

    return J.Model(prior_model, J.MultivariateNormalLikelihood(mu, cov=Sig)), float(log_Z), mu, Sig


CONFIGS = {
    "c2": dict(D=32, N=2048, s=4, name="MVN D=32, num_live_points=2048, s=4 (FP32 likelihood, FP64 evidence)"),
    "c1": dict(D=10, N=300, s=5, name="MVN D=10, num_live_points=300, s=5 (FP32 likelihood, FP64 evidence)"),
}


def make_batch(model, cfg, seeds, lpc, wpb):
    from jaxns_amd.engine import DeviceBatch, RunSpec
    D, N, s = cfg["D"], cfg["N"], cfg["s"]
    specs = [RunSpec(model=model, num_live_points=N, num_slices=D * s, k_phantom=0, max_samples=N * 120, seed=sd)
             for sd in seeds]
    return DeviceBatch(specs, lanes_per_chain=lpc, waves_per_block=wpb, max_steps=512)


def term_for(batch):
    return dict(dlogZ=math.log(1.0 + 1e-3), max_samples=float(batch.specs[0].max_samples))


def run_batch(batch, graph_iters):
    batch.run(term_for(batch), init_efficiency_threshold=0.1, graph_iters=graph_iters)
    return batch.log_Z()


def instrumented_slice_timing(model, cfg, R, lpc, wpb, seed0):
    """Runs one batch shell by shell with HIP events around the sampling kernel
    (hipns_shell_sample) on the launch stream; returns per-launch durations and
    the likelihood evaluations of each launch."""
    import torch
    from jaxns_amd._lib import check, lib, ptr, stream_ptr
    b = make_batch(model, cfg, range(seed0, seed0 + R), lpc, wpb)
    term = term_for(b)
    b.init_state()
    b.check_termination(dict(efficiency_threshold=0.1, dlogZ=0.0, max_samples=term["max_samples"]), True)
    b.check_termination(term, True)
    tc = __import__("jaxns_amd._lib", fromlist=["make_term_cond"]).make_term_cond(**term)
    durs, evals = [], []
    s = stream_ptr()
    prev = sum(st.num_likelihood_evaluations for st in b.states())
    while not np.all(b.done_flags()):
        check(lib().hipns_shell_discard(ptr(b.runs_dev), ptr(b.states_dev), b.R, b.runs_host, C.byref(b.bufs), s))
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        check(lib().hipns_shell_sample(ptr(b.runs_dev), ptr(b.states_dev), b.R, b.runs_host, C.byref(b.bufs),
                                       C.byref(b.launch), s))
        e1.record()
        check(lib().hipns_shell_replace(ptr(b.runs_dev), ptr(b.states_dev), b.R, b.runs_host, C.byref(b.bufs), s))
        check(lib().hipns_shell_register(ptr(b.runs_dev), ptr(b.states_dev), b.R, b.runs_host, C.byref(b.bufs),
                                         C.byref(tc), s))
        torch.cuda.synchronize()
        cur = sum(st.num_likelihood_evaluations for st in b.states())
        durs.append(e0.elapsed_time(e1) * 1e-3)
        evals.append(cur - prev)
        prev = cur
    return np.array(durs), np.array(evals)


def cpu_baseline(mu, Sig, cfg, gpu_evals_per_run, budget_s=15.0):
    """The oracle (a C port of the reference algorithm) on the host cores for a bounded
    number of shells of the same workload; evals/s extrapolated to runs/s with the
    evaluations a full run needs."""
    import oracle as O
    D = cfg["D"]
    spec = O.ModelSpec(O.FAM_MVN, D, O.PRIOR_NORMAL, np.zeros(D), np.ones(D), mvn_mu=mu,
                       mvn_chol=np.linalg.cholesky(Sig))
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    iters = 2
    while True:
        t0 = time.perf_counter()
        r = O.ns_run(spec, num_live_points=cfg["N"], num_slices=D * cfg["s"], max_samples=cfg["N"] * 120, seed=5,
                     num_threads=threads, max_iterations=iters)
        dt = time.perf_counter() - t0
        print(f"[bench] cpu baseline: {iters} shells in {dt:.2f} s", file=sys.stderr, flush=True)
        if dt > budget_s / 3 or iters >= 64:
            break
        iters *= 2
    evals = r["total_likelihood_evaluations"]
    eps = evals / dt
    return dict(value=eps / gpu_evals_per_run, unit="runs/s", cores=threads, kind="port",
                evals_per_s=eps,
                sample=f"{iters} shells of one run of the same workload (create_init_state + {iters} x "
                       f"_collect_shell + to_results), {evals} likelihood evals in {dt:.2f} s, "
                       f"extrapolated with {gpu_evals_per_run:.3g} evals per full run")


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--runs", type=int, default=128, help="runs per step (chosen form)")
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--graph-iters", type=int, default=8)
    ap.add_argument("--lpc", type=int, default=4)
    ap.add_argument("--wpb", type=int, default=4)
    ap.add_argument("--plain-runs", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="use --lpc/--wpb instead of the autotuner")
    ap.add_argument("--profile-pass", action="store_true",
                    help="only the instrumented shell-by-shell pass (for rocprofv3 kernel statistics)")
    ap.add_argument("--copy-calib", type=int, default=0,
                    help="with --profile-pass: also copy a buffer of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box in this build has one GPU; "
              "multi-GPU runs would shard independent runs across processes", file=sys.stderr)
        sys.exit(2)
    import torch
    torch.cuda.set_device(0)
    cfg = CONFIGS[args.config]
    model, log_Z_true, mu, Sig = mvn_model(cfg["D"], seed=0)
    R = args.runs
    if args.profile_pass:
        if args.copy_calib:
            n = args.copy_calib * (1 << 20) // 4
            a = torch.ones(n, dtype=torch.float32, device="cuda")
            b = torch.empty_like(a)
            for _ in range(3):
                b.copy_(a)
            torch.cuda.synchronize()
            log(f"copy calibration: 3 x {args.copy_calib} MiB read + written")
        durs, evs = instrumented_slice_timing(model, cfg, R, args.lpc, args.wpb, seed0=90000)
        D = cfg["D"]
        print(json.dumps({"profile_pass": True, "runs": R, "lanes_per_chain": args.lpc, "waves_per_block": args.wpb,
                          "sample_launches": int(durs.size), "avg_launch_ms": float(durs.mean() * 1e3),
                          "evals_per_launch": float(evs.mean()),
                          "achieved_TFLOPs": float((evs * 2.0 * D * D).mean() / durs.mean() / 1e12)}), flush=True)
        return
    if not args.no_autotune:
        from jaxns_amd.autotune import autotune
        t0 = time.perf_counter()
        args.lpc, args.wpb = autotune(model, cfg["N"], cfg["D"] * cfg["s"], R, verbose=True)
        log(f"autotune picked lanes_per_chain={args.lpc} waves_per_block={args.wpb} "
            f"in {time.perf_counter() - t0:.1f} s")

    # ---------------- chosen form: R runs per step
    for w in range(args.warmup):
        b = make_batch(model, cfg, range(1000 + w * R, 1000 + (w + 1) * R), args.lpc, args.wpb)
        t0 = time.perf_counter()
        run_batch(b, args.graph_iters)
        log(f"warmup {w}: {R} runs in {time.perf_counter() - t0:.2f} s")
        del b
    torch.cuda.synchronize()
    logZ = []
    evals_total = 0
    t_total = 0.0
    for k in range(args.steps):
        b = make_batch(model, cfg, range(k * R, (k + 1) * R), args.lpc, args.wpb)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.run(term_for(b), init_efficiency_threshold=0.1, graph_iters=args.graph_iters)
        t1 = time.perf_counter()
        z = b.log_Z()
        torch.cuda.synchronize()
        t_total += time.perf_counter() - t0
        log(f"step {k}: sampling {t1 - t0:.3f} s, log Z of {R} runs {time.perf_counter() - t1:.3f} s")
        logZ.append(z)
        evals_total += sum(st.num_likelihood_evaluations for st in b.states())
        log(f"step {k}: {R} runs, total {t_total:.2f} s, mean logZ {z[:, 0].mean():.3f} (true {log_Z_true:.3f})")
        del b
    logZ = np.concatenate(logZ)
    ms_per_step = 1e3 * t_total / args.steps
    runs_per_s = R * args.steps / t_total
    evals_per_s = evals_total / t_total
    dz = logZ[:, 0] - log_Z_true
    evals_per_run = evals_total / (R * args.steps)

    # ---------------- plain form: one run at a time, host sync per shell
    tp = 0.0
    pz = []
    for k in range(args.plain_runs):
        b = make_batch(model, cfg, [50000 + k], args.lpc, args.wpb)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pz.append(run_batch(b, 0)[0])
        torch.cuda.synchronize()
        tp += time.perf_counter() - t0
        log(f"plain run {k}: {tp:.2f} s so far")
        del b
    plain_runs_per_s = args.plain_runs / tp

    # ---------------- roofline of the dominant kernel (hipns_shell_sample)
    durs, evs = instrumented_slice_timing(model, cfg, R, args.lpc, args.wpb, seed0=90000)
    log(f"instrumented: {durs.size} sample launches, mean {durs.mean() * 1e3:.3f} ms")
    D = cfg["D"]
    flops = evs * 2.0 * D * D  # MVN residual GEMM: 2 D^2 flops per likelihood evaluation
    avg_dur = float(durs.mean())
    achieved = float(flops.mean() / avg_dur)
    roofline = dict(bound="mfma", achieved=achieved / 1e12, peak=PEAK_FP32_MFMA / 1e12, unit="TFLOP/s",
                    frac=achieved / PEAK_FP32_MFMA, traffic=None,
                    kernel="k_sample<slice, MVN> (hipns_shell_sample)", avg_launch_ms=avg_dur * 1e3,
                    launches=int(durs.size), evals_per_launch=float(evs.mean()),
                    evals_per_s_in_kernel=float(evs.sum() / durs.sum()))

    cpu = None if args.no_cpu else cpu_baseline(mu, Sig, cfg, evals_per_run)
    log("cpu baseline", cpu)
    out = {
        "metric": "runs/s to |dlogZ|<0.1 nat (chosen batched form, 1 GPU)",
        "value": runs_per_s,
        "unit": "runs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 likelihood (MFMA) + f64 evidence",
        "data": "synthetic (seeded numpy)",
        "config": {"workload": cfg["name"], "runs_per_step": R, "graph_iters": args.graph_iters,
                   "lanes_per_chain": args.lpc, "waves_per_block": args.wpb},
        "likelihood_evals_per_s": evals_per_s,
        "evals_per_run": evals_per_run,
        "plain_form_runs_per_s": plain_runs_per_s,
        "chosen_vs_plain": runs_per_s / plain_runs_per_s,
        "log_Z_true": log_Z_true,
        "mean_dlogZ": float(dz.mean()),
        "mean_abs_dlogZ": float(np.abs(dz).mean()),
        "mean_stated_uncert": float(logZ[:, 1].mean()),
        "frac_z_lt_3": float(np.mean(np.abs(dz) / logZ[:, 1] < 3)),
        "ensemble_std_logZ": float(logZ[:, 0].std()),
        "plain_dlogZ": [float(z[0] - log_Z_true) for z in pz],
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
