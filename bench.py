"""Benchmark: runs/s to |dlogZ| < 0.1 nat and likelihood evals/s on one MI355X.

Workload (default, --config c2): D=32 multivariate-normal likelihood, prior
N(0, I), seeded mu ~ N(0, 0.5^2 I), Sigma = A A^T / D + 0.05 I;
num_live_points=2048, s=4 (128 slices per chain, 1024 chains per shell); FP32
likelihood (MFMA), FP64 evidence.  One step = one batch of R independent runs
(different seeds) taken from create_init_state to log Z, in the chosen form:
all R runs advanced by the same kernel launches, shells replayed from a
captured hipGraph.  The plain form (one run at a time, host sync after every
shell) is timed beside it.

Other workloads:
  c1  MVN D=10, num_live_points=300, s=5
  c3  runs of uneven D in {4..64} and num_live_points in {128..1024} (1024 per step by
      default), 4-component isotropic Gaussian mixture, prior U[-10, 10]^D
  c4  Bayesian linear regression, 4096 rows x 16 features, sigma=0.3,
      w ~ N(0, I), num_live_points=1024, s=5; residual GEMM in BF16x3 MFMA

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

# HBM bytes per k_sample launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on the same
# instrumented batch (counters scaled by a 256 MiB copy: FETCH_SIZE x 2.0, WRITE_SIZE x 1.0); see
# profiles/k_sample_mvn_d32_n2048_r1024.md and profiles/k_sample_linreg_4096x16_n1024_r128.md.
# Keyed by (workload, runs per launch); other shapes report null.
MEASURED_TRAFFIC = {("c2", 1024): 370.0e6, ("c4", 128): 17.68e6, ("c4", 256): 32.55e6}
PEAK_BF16_MFMA = 2.5e15
PEAK_HBM = 8.0e12
# wavefront VALU instructions per second: 256 CUs x 4 SIMDs x 2.4 GHz, one wave64 instruction per 4
# cycles per SIMD (16 lanes).  The 157.3 TFLOP/s FP32 vector peak counts packed FP32 (two FMAs per lane
# per instruction), which these kernels are built without; the counters agree: SQ_ACTIVE_INST_VALU x 4
# cycles per SQ_INSTS_VALU is 4.3 on the MVN sampling kernel
PEAK_VALU_ISSUE = 256 * 4 * 2.4e9 / 4.0
# SQ_INSTS_VALU per k_sample launch (rocprofv3 --pmc on the bench's instrumented pass of the same
# shape), keyed like MEASURED_TRAFFIC
MEASURED_VALU = {("c2", 1024): 1.1726e10}
# SQ_INSTS_VALU per likelihood evaluation of the sampling kernel (rocprofv3 --pmc on the instrumented
# pass: mixture batch of 256 runs at 8 lanes per chain and one wavefront per workgroup, MVN D=10 batch of
# 1024 runs at 2 lanes per chain and one wavefront per workgroup), keyed by (workload, lanes per chain);
# times a step's evaluations per second it gives the VALU issue rate the step sustains
# (profiles/k_sample_gmm_mixture_r256.md, profiles/k_sample_mvn_d10_n300_r1024.md)
MEASURED_VALU_PER_EVAL = {("c3", 8): 82.58, ("c1", 2): 21.74}
# the oracle ensemble of the same workload (tests/golden/oracle_logZ_ensembles.json) beside the
# device's log Z statistics
# runs of a step streamed through this many device slots (engine.run_queue) instead of one batch: none
# by default (mixture batch, one step: 1024 runs in one batch 76.4 runs/s, 2048 in one batch 83.6,
# 2048 through 1024 slots 78.4; a larger launch amortises each shell's slowest block better than slot
# refill shortens the end of the step)
QUEUE_SLOTS = {}
ORACLE_ENSEMBLE_KEY = {"c1": "mvn_d10", "c2": "mvn_d32", "c4": "linreg", "c3": "gmm"}


def oracle_ensemble(name):
    """Summary of the oracle's ensemble of this workload from the committed fixture (or None)."""
    path = os.path.join(ROOT, "tests", "golden", "oracle_logZ_ensembles.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return dict(d[name]["summary"], source="tests/golden/oracle_logZ_ensembles.json")
    except (OSError, KeyError, ValueError):
        return None


# ----------------------------------------------------------------------------- workloads
class Workload:
    """Runs of one benchmark configuration: specs per seed, analytic log Z per run, the
    algorithmic work of one likelihood evaluation, and the oracle model for the CPU leg."""

    name = ""
    dtype = ""
    kernel = ""
    bound = "mfma"
    unit = "TFLOP/s"
    peak = PEAK_FP32_MFMA
    lr_terms = 3

    def specs(self, seeds):
        raise NotImplementedError

    def work_per_eval(self, spec) -> float:
        raise NotImplementedError

    def oracle_spec(self, spec):
        raise NotImplementedError


class MvnWorkload(Workload):
    kernel = "k_sample<slice, MVN> (hipns_shell_sample)"
    dtype = "f32 likelihood (MFMA) + f64 evidence"

    def __init__(self, D, N, s):
        from problems import mvn_problem
        self.D, self.N, self.s = D, N, s
        self.name = f"MVN D={D}, num_live_points={N}, s={s} (FP32 likelihood, FP64 evidence)"
        self.model, self.log_Z, self.mu, self.Sig = mvn_problem(D, seed=0)

    def specs(self, seeds):
        from jaxns_amd.engine import RunSpec
        sp = [RunSpec(model=self.model, num_live_points=self.N, num_slices=self.D * self.s, k_phantom=0,
                      max_samples=self.N * 120, seed=sd) for sd in seeds]
        return sp, np.full(len(sp), self.log_Z)

    def work_per_eval(self, spec):
        return 2.0 * self.D * self.D  # residual GEMM W z: 2 D^2 flops

    def oracle_spec(self, spec):
        from problems import mvn_oracle_spec
        return mvn_oracle_spec(self.D, self.mu, self.Sig)


class GmmBatchWorkload(Workload):
    """Run r draws D in {4..64} and num_live_points in {128..1024} from a generator seeded by r."""
    kernel = "k_sample<slice, GMM> (hipns_shell_sample)"
    dtype = "f32 likelihood + f64 evidence"
    bound = "hbm"
    unit = "GB/s"
    peak = PEAK_HBM

    def __init__(self):
        self.name = ("runs with D~U{4..64}, num_live_points~U{128..1024}, 4-component isotropic Gaussian mixture, "
                     "prior U[-10,10]^D, s=5 (FP32 likelihood, FP64 evidence)")
        self._cache = {}

    def _run(self, seed):
        from problems import gmm_problem
        if seed not in self._cache:
            rng = np.random.default_rng(seed)
            D = int(rng.integers(4, 65))
            N = int(rng.integers(128, 1025))
            model, log_Z, means, sig = gmm_problem(D, seed=10_000 + seed)
            self._cache[seed] = (D, N, model, log_Z, means, sig)
        return self._cache[seed]

    def specs(self, seeds):
        from jaxns_amd.engine import RunSpec
        sp, tr = [], []
        for sd in seeds:
            D, N, model, log_Z, _, _ = self._run(sd)
            sp.append(RunSpec(model=model, num_live_points=N, num_slices=5 * D, k_phantom=0, max_samples=N * 300,
                              seed=sd))
            tr.append(log_Z)
        return sp, np.array(tr)

    def work_per_eval(self, spec):
        return 4.0 * spec.model.U_ndims + 4.0  # read the proposal (FP32) + write logL: algorithmic bytes

    def oracle_spec(self, spec):
        from problems import gmm_oracle_spec
        for D, N, model, log_Z, means, sig in self._cache.values():
            if model is spec.model:
                return gmm_oracle_spec(D, means, sig)
        raise KeyError("unknown run")


class LinRegWorkload(Workload):
    kernel = "k_sample<slice, LinReg BF16x3> (hipns_shell_sample)"
    dtype = "bf16x3 likelihood GEMM (MFMA, f32 accumulate) + f64 evidence"
    peak = PEAK_BF16_MFMA

    def __init__(self, Ndata=4096, D=16, N=1024, s=5, sigma=0.3, tau=1.0, lr_terms=3):
        from problems import linreg_problem
        self.Ndata, self.D, self.N, self.s, self.tau = Ndata, D, N, s, tau
        self.lr_terms = lr_terms
        self.name = (f"Bayesian linear regression {Ndata} x {D}, sigma={sigma}, w ~ N(0, {tau}^2 I), "
                     f"num_live_points={N}, s={s} (BF16x{lr_terms} likelihood GEMM, FP64 evidence)")
        self.model, self.log_Z, self.lik = linreg_problem(Ndata, D, seed=0, sigma=sigma, tau=tau)

    def specs(self, seeds):
        from jaxns_amd.engine import RunSpec
        sp = [RunSpec(model=self.model, num_live_points=self.N, num_slices=self.D * self.s, k_phantom=0,
                      max_samples=self.N * 200, seed=sd) for sd in seeds]
        return sp, np.full(len(sp), self.log_Z)

    def work_per_eval(self, spec):
        return 2.0 * self.Ndata * self.D + 2.0 * self.Ndata  # X dw and the squared residual sum

    def oracle_spec(self, spec):
        from problems import linreg_oracle_spec
        return linreg_oracle_spec(self.lik, self.tau)


class LinRegScaleWorkload(LinRegWorkload):
    """Precision at scale: 65536 rows x 128 features, 10 % of the columns x 100, sigma=0.1,
    w ~ N(0, I); num_live_points=4096, s=3."""

    def __init__(self):
        from problems import linreg_problem
        self.Ndata, self.D, self.N, self.s, self.tau = 65536, 128, 4096, 3, 1.0
        self.lr_terms = 3
        self.name = ("Bayesian linear regression 65536 x 128 (10% of columns x100), sigma=0.1, w ~ N(0, I), "
                     "num_live_points=4096, s=3 (BF16x3 likelihood GEMM, FP64 evidence)")
        self.model, self.log_Z, self.lik = linreg_problem(65536, 128, seed=0, sigma=0.1, tau=1.0, scaled_cols=0.1)


def make_workload(name):
    if name == "c1":
        return MvnWorkload(10, 300, 5)
    if name == "c2":
        return MvnWorkload(32, 2048, 4)
    if name == "c3":
        return GmmBatchWorkload()
    if name == "c4":
        return LinRegWorkload()
    if name == "c5":
        return LinRegScaleWorkload()
    raise ValueError(name)


def make_batch(wl, seeds, lpc, wpb, plain=False):
    """plain=True: the plain GPU form's arithmetic (FP32 GEMM for the regression likelihood)."""
    from jaxns_amd.engine import DeviceBatch
    specs, truths = wl.specs(seeds)
    terms = 1 if (plain and isinstance(wl, LinRegWorkload)) else wl.lr_terms
    return DeviceBatch(specs, lanes_per_chain=lpc, waves_per_block=wpb, max_steps=512, lr_terms=terms), truths


def term_for(batch):
    return dict(dlogZ=math.log(1.0 + 1e-3), max_samples=float(max(s.max_samples for s in batch.specs)))


def run_batch(batch, graph_iters):
    batch.run(term_for(batch), init_efficiency_threshold=0.1, graph_iters=graph_iters)
    return batch.log_Z()


def make_groups(wl, seeds, lpc, wpb, groups):
    """The runs of one step split into `groups` batches (contiguous seed ranges)."""
    seeds = list(seeds)
    cut = np.linspace(0, len(seeds), groups + 1).round().astype(int)
    out = [make_batch(wl, seeds[cut[g]:cut[g + 1]], lpc, wpb) for g in range(groups)]
    return [b for b, _ in out], np.concatenate([t for _, t in out])


def run_groups(batches, graph_iters):
    """The chosen form with several groups: every group on its own stream, shells of one group running
    while another's drain (engine.run_concurrent); log Z of every run, in seed order."""
    from jaxns_amd.engine import run_concurrent
    if len(batches) == 1:
        batches[0].run(term_for(batches[0]), init_efficiency_threshold=0.1, graph_iters=graph_iters)
    else:
        run_concurrent(batches, [term_for(b) for b in batches], init_efficiency_threshold=0.1,
                       graph_iters=graph_iters)


def instrumented_slice_timing(wl, R, lpc, wpb, seed0):
    """Runs one batch shell by shell with HIP events around the sampling kernel
    (hipns_shell_sample) on the launch stream; returns per-launch durations and
    the algorithmic work (flops or bytes) and likelihood evaluations of each launch."""
    import torch
    from jaxns_amd._lib import check, lib, make_term_cond, ptr, stream_ptr
    b, _ = make_batch(wl, range(seed0, seed0 + R), lpc, wpb)
    term = term_for(b)
    b.init_state()
    b.check_termination(dict(efficiency_threshold=0.1, dlogZ=0.0, max_samples=term["max_samples"]), True)
    b.check_termination(term, True)
    tc = make_term_cond(**term)
    wpe = np.array([wl.work_per_eval(sp) for sp in b.specs])
    durs, work, evals, util = [], [], [], []
    CW = 64 // (4 if b.dps > 64 else lpc)
    s = stream_ptr()
    prev = np.array([st.num_likelihood_evaluations for st in b.states()], dtype=np.float64)
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
        cur = np.array([st.num_likelihood_evaluations for st in b.states()], dtype=np.float64)
        # lane utilisation of the launch: chain steps over wavefront steps (a wavefront runs until
        # its slowest chain finishes)
        ne = b.new_evals.cpu().numpy().astype(np.float64)
        busy = tot = 0.0
        for r, sp in enumerate(b.specs):
            if cur[r] == prev[r]:
                continue
            e = ne[b.runs_host[r].chain_off:b.runs_host[r].chain_off + sp.shell]
            pad = (-e.size) % CW
            e = np.concatenate([e, np.zeros(pad)]).reshape(-1, CW)
            busy += e.sum()
            tot += e.max(axis=1).sum() * CW
        util.append(busy / max(tot, 1.0))
        durs.append(e0.elapsed_time(e1) * 1e-3)
        work.append(float(((cur - prev) * wpe).sum()))
        evals.append(float((cur - prev).sum()))
        prev = cur
    instrumented_slice_timing.lane_util = float(np.mean(util)) if util else float("nan")
    b.log_Z()  # the batched _to_results evidence of the same runs (its kernels show in a profile)
    return np.array(durs), np.array(work), np.array(evals)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(wl, spec, gpu_evals_per_run, budget_s=12.0):
    """The oracle (a C port of the reference algorithm) on the host cores.  A shell-limited probe
    measures evaluations/s; when a whole run is predicted to fit twice in `budget_s`, whole runs of
    the same workload (different seeds) are timed and runs/s is reported directly, otherwise the
    probe is grown to about a third of the budget and extrapolated with the device's evaluations
    per run.  A shell-limited sample on one thread gives the single-core rate beside it."""
    import oracle as O
    ospec = wl.oracle_spec(spec)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))

    def run(seed, iters, nthreads=threads):
        t0 = time.perf_counter()
        r = O.ns_run(ospec, num_live_points=spec.num_live_points, num_slices=spec.num_slices,
                     max_samples=spec.max_samples, seed=seed, num_threads=nthreads, max_iterations=iters)
        return time.perf_counter() - t0, r["total_likelihood_evaluations"]

    def one_core():
        it = 1
        while True:
            dt1, ev1 = run(5, it, 1)
            if dt1 > 1.5 or it >= 256:
                break
            it *= 2
        print(f"[bench] cpu baseline, one thread: {it} shells in {dt1:.2f} s", file=sys.stderr, flush=True)
        return dict(evals_per_s=ev1 / dt1, runs_per_s=ev1 / dt1 / gpu_evals_per_run,
                    sample=f"{it} shells of one run on one thread, {ev1} likelihood evaluations in {dt1:.2f} s, "
                           f"runs/s extrapolated with {gpu_evals_per_run:.3g} evaluations per full run")

    info = dict(cpu_model=cpu_model(), nproc=os.cpu_count())
    iters = 2
    while True:
        dt, ev = run(5, iters)
        print(f"[bench] cpu baseline probe: {iters} shells in {dt:.2f} s", file=sys.stderr, flush=True)
        if dt > 0.3 or iters >= 64:
            break
        iters *= 2
    est_run = gpu_evals_per_run / (ev / dt)
    if 2.0 * est_run < budget_s:
        t_all, evals, runs = 0.0, 0, 0
        while t_all < budget_s and runs < 32:
            dt, ev = run(11 + runs, -1)
            t_all += dt
            evals += ev
            runs += 1
            print(f"[bench] cpu baseline: run {runs} in {dt:.2f} s", file=sys.stderr, flush=True)
        return dict(value=runs / t_all, unit="runs/s", cores=threads, kind="port", evals_per_s=evals / t_all,
                    sample=f"{runs} whole runs of the same workload (create_init_state, shells until the "
                           f"termination condition, to_results; D={spec.model.U_ndims}, num_live_points="
                           f"{spec.num_live_points}, different seeds), {evals} likelihood evaluations in "
                           f"{t_all:.2f} s", one_core=one_core(), **info)
    while dt < budget_s / 3 and iters < 4096:
        iters *= 2
        dt, ev = run(5, iters)
        print(f"[bench] cpu baseline probe: {iters} shells in {dt:.2f} s", file=sys.stderr, flush=True)
    eps = ev / dt
    return dict(value=eps / gpu_evals_per_run, unit="runs/s", cores=threads, kind="port", evals_per_s=eps,
                sample=f"{iters} shells of one run of the same workload (create_init_state + {iters} x "
                       f"_collect_shell + to_results; D={spec.model.U_ndims}, num_live_points="
                       f"{spec.num_live_points}), {ev} likelihood evaluations in {dt:.2f} s, extrapolated "
                       f"with {gpu_evals_per_run:.3g} evaluations per full run", one_core=one_core(), **info)


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def precision_at_scale(wl, R, shells, lpc, wpb, steps):
    """Partial runs of the 65536 x 128 regression (a full run needs ~1e9 likelihood evaluations):
    `shells` shells of R runs in each product form, timed with HIP events around the sampling kernel,
    plus the logL error of the accepted points against the FP64 likelihood on the host."""
    import torch
    from jaxns_amd._lib import check, lib, make_term_cond, ptr, stream_ptr
    from jaxns_amd.engine import DeviceBatch, RunSpec
    out = {}
    for terms, label in ((3, "bf16x3"), (1, "fp32")):
        specs = [RunSpec(model=wl.model, num_live_points=wl.N, num_slices=wl.D * wl.s, k_phantom=0,
                         max_samples=wl.N * (shells + 4), seed=sd) for sd in range(R)]
        refine = terms != 1
        b = DeviceBatch(specs, lanes_per_chain=lpc, waves_per_block=wpb, max_steps=512, lr_terms=terms,
                        refine_fp32=refine)
        b.init_state()
        tc = make_term_cond(max_samples=float(specs[0].max_samples))
        s = stream_ptr()
        durs, evs, rdurs = [], [], []
        prev = sum(st.num_likelihood_evaluations for st in b.states())
        errs, raw_errs = [], []
        for k in range(shells):
            check(lib().hipns_shell_discard(ptr(b.runs_dev), ptr(b.states_dev), b.R, b.runs_host, C.byref(b.bufs), s))
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            check(lib().hipns_shell_sample(ptr(b.runs_dev), ptr(b.states_dev), b.R, b.runs_host, C.byref(b.bufs),
                                           C.byref(b.launch), s))
            e1.record()
            torch.cuda.synchronize()
            # accepted points of run 0 against the FP64 likelihood (prior transform in FP64 as well)
            n0 = specs[0].shell
            U = b.new_U[:n0, :wl.D].double().cpu().numpy()
            Ld = b.new_logL[:n0].cpu().numpy()
            sel = np.linspace(0, n0 - 1, 64).astype(int)
            from scipy.special import ndtri
            Lh = wl.lik.log_prob_numpy(wl.tau * ndtri(U[sel]))
            if refine:
                # the accepted points re-evaluated in plain FP32 before they enter the live set
                raw_errs.append(np.abs(Ld[sel] - Lh) / np.abs(Lh))
                r0, r1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                r0.record()
                b.refine_new()
                r1.record()
                torch.cuda.synchronize()
                rdurs.append(r0.elapsed_time(r1) * 1e-3)
                Ld = b.new_logL[:n0].cpu().numpy()
            errs.append(np.abs(Ld[sel] - Lh) / np.abs(Lh))
            check(lib().hipns_shell_replace(ptr(b.runs_dev), ptr(b.states_dev), b.R, b.runs_host, C.byref(b.bufs), s))
            check(lib().hipns_shell_register(ptr(b.runs_dev), ptr(b.states_dev), b.R, b.runs_host,
                                             C.byref(b.bufs), C.byref(tc), s))
            torch.cuda.synchronize()
            cur = sum(st.num_likelihood_evaluations for st in b.states())
            durs.append(e0.elapsed_time(e1) * 1e-3)
            evs.append(cur - prev)
            prev = cur
            log(f"{label} shell {k}: {durs[-1] * 1e3:.1f} ms, {evs[-1]} evals")
        durs, evs = np.array(durs), np.array(evs, dtype=np.float64)
        eps = float(evs.sum() / durs.sum())
        e = np.concatenate(errs)
        out[label] = dict(evals_per_s=eps, avg_launch_ms=float(durs.mean() * 1e3), evals_per_launch=float(evs.mean()),
                          logL_rel_err_max=float(e.max()), logL_rel_err_median=float(np.median(e)),
                          achieved_TFLOPs=eps * wl.work_per_eval(None) / 1e12)
        if refine:
            er = np.concatenate(raw_errs)
            out[label].update(refined_fp32=True, refine_ms_per_shell=float(np.mean(rdurs) * 1e3),
                              unrefined_logL_rel_err_max=float(er.max()),
                              unrefined_logL_rel_err_median=float(np.median(er)))
        del b
    return out


def full_runs_at_scale(wl, R, wpb, graph_iters):
    """R whole runs of the 65 536 x 128 regression to termination in one batch (BF16x3 acceptance, accepted
    points refined in plain FP32 before they enter the live set), one chain group per workgroup so that
    even a few runs spread over every CU; log Z of each run against the analytic evidence."""
    import torch
    from jaxns_amd.engine import DeviceBatch, RunSpec
    specs = [RunSpec(model=wl.model, num_live_points=wl.N, num_slices=wl.D * wl.s, k_phantom=0,
                     max_samples=wl.N * 1600, seed=70_000 + i) for i in range(R)]
    b = DeviceBatch(specs, lanes_per_chain=4, waves_per_block=wpb, max_steps=512, lr_terms=3, refine_fp32=True)
    term = dict(dlogZ=math.log(1.0 + 1e-3), max_samples=float(specs[0].max_samples))
    t0 = time.perf_counter()
    b.init_state()
    b.check_termination(dict(efficiency_threshold=0.1, dlogZ=0.0, max_samples=term["max_samples"]), True)
    b.check_termination(term, True)
    shells = 0
    while not np.all(b.done_flags()):
        for _ in range(graph_iters):
            b.collect_shell(term)
        shells += graph_iters
        torch.cuda.synchronize()
        st = b.states()
        log(f"c5 full runs: {shells} shells, {time.perf_counter() - t0:.0f} s, contour {st[0].log_L_contour:.6g}, "
            f"evals {sum(x.num_likelihood_evaluations for x in st):.4g}")
    b.consume()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    z = b.log_Z()
    sts = b.states()
    evals = sum(x.num_likelihood_evaluations for x in sts)
    dz = z[:, 0] - wl.log_Z
    return dict(runs=R, wall_s=wall, runs_per_s=R / wall, evals_per_s=evals / wall, evals_per_run=evals / R,
                shells=[int(x.iteration) for x in sts], log_Z=[float(v) for v in z[:, 0]],
                log_Z_uncert=[float(v) for v in z[:, 1]], log_Z_true=float(wl.log_Z), dlogZ=[float(v) for v in dz],
                z_scores=[float(v) for v in dz / z[:, 1]], cap_hits=int(b.cap_hits().sum()),
                refined_below_contour=[int(v) for v in b.refine_below.cpu().numpy()],
                termination_reason=[int(x.termination_reason) for x in sts], waves_per_block=wpb)


def to_results_scale(M, iters=5):
    """The post-path primitives of _to_results at M samples, each timed with HIP events on the stream it
    runs on: the evidence scan (compute_evidence_stats, 80 algorithmic bytes per sample: read log L and n,
    write the 8-field EvidenceCalculation), count_crossed_edges (28 B per node: read sender and log L,
    write the sample index and the live-point count) and the stable FP64 argsort (16 B per key)."""
    import torch
    from jaxns_amd._lib import check, lib, ptr, stream_ptr
    L = lib()
    g = torch.Generator().manual_seed(11)
    keys = torch.randn(M, dtype=torch.float64, generator=g).cuda()
    logL_sorted = torch.sort(keys)[0]
    # a static run's live-point counts: N throughout, then N..1 over the last N samples
    N = 4096
    n = torch.full((M,), float(N), dtype=torch.float64, device="cuda")
    n[M - N:] = torch.arange(N, 0, -1, dtype=torch.float64, device="cuda")
    # sender of sample i: a uniform earlier node (0 is the root)
    sender = (torch.rand(M, generator=g, dtype=torch.float64) * torch.arange(1, M + 1, dtype=torch.float64))
    sender = sender.floor().to(torch.int32).cuda()
    per = torch.empty((8, M), dtype=torch.float64, device="cuda")
    fin = torch.empty(8, dtype=torch.float64, device="cuda")
    si = torch.empty(M, dtype=torch.int64, device="cuda")
    nl = torch.empty(M, dtype=torch.int64, device="cuda")
    perm = torch.empty(M, dtype=torch.int64, device="cuda")
    ws_e = torch.empty(int(L.hipns_evidence_scan_workspace_size(M)), dtype=torch.uint8, device="cuda")
    ws_c = torch.empty(int(L.hipns_count_crossed_edges_workspace_size(M)), dtype=torch.uint8, device="cuda")
    ws_a = torch.empty(int(L.hipns_argsort_workspace_size(M)), dtype=torch.uint8, device="cuda")
    s = stream_ptr()
    calls = {
        "evidence_scan": (80, lambda: check(L.hipns_evidence_scan(ptr(logL_sorted), ptr(n), M, -1, ptr(per), ptr(fin),
                                                                  ptr(ws_e), ws_e.numel(), s), "hipns_evidence_scan")),
        "count_crossed_edges": (28, lambda: check(L.hipns_count_crossed_edges(ptr(sender), ptr(keys), M, -1, ptr(si),
                                                                              ptr(nl), ptr(ws_c), ws_c.numel(), s),
                                                  "hipns_count_crossed_edges")),
        "argsort_stable": (16, lambda: check(L.hipns_argsort_stable(ptr(keys), M, ptr(perm), ptr(ws_a), ws_a.numel(), s),
                                             "hipns_argsort_stable")),
    }
    out = {}
    for name, (bps, fn) in calls.items():
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(iters):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / iters
        gbs = bps * M / (ms * 1e-3) / 1e9
        out[name] = dict(avg_ms=ms, algorithmic_bytes=bps * M, achieved_GBps=gbs, frac=gbs * 1e9 / PEAK_HBM)
        log(name, f"{ms:.3f} ms, {gbs:.0f} GB/s")
    # the argsort agrees with torch's stable sort (the sort of count_crossed_edges is the same code)
    ok = bool(torch.equal(perm, torch.sort(keys, stable=True)[1]))
    return dict(to_results_scale=True, M=M, iters=iters, kernels=out, argsort_matches_torch=ok)


def run_step_queue(wl, seeds, lpc, wpb, slots, graph_iters):
    """One step of the chosen form through engine.run_queue: the runs stream through `slots` device slots;
    returns ([R, 2] log Z, truths, evaluations, step-cap hits)."""
    from jaxns_amd.engine import run_queue
    specs, truths = wl.specs(seeds)
    res = run_queue(specs, slots, dict(dlogZ=math.log(1.0 + 1e-3), max_samples=float(max(s.max_samples for s in specs))),
                    graph_iters=graph_iters, lanes_per_chain=lpc, waves_per_block=wpb, max_steps=512,
                    lr_terms=wl.lr_terms)
    z = np.array([(r["log_Z_mean"], r["log_Z_uncert"]) for r in res])
    return z, truths, sum(r["num_likelihood_evaluations"] for r in res), sum(r["cap_hits"] for r in res)


def measure_chosen(wl, cfg, R, G, steps, warmup, graph_iters, lpc, wpb, tune, seed_base=0, slots=0):
    """The chosen form on one workload: `warmup` untimed then `steps` timed steps of R runs (G concurrent
    groups, or streamed through `slots` slots when slots > 0), each from create_init_state to log Z.
    Returns the step statistics."""
    import torch
    if tune:
        from jaxns_amd.autotune import autotune
        t0 = time.perf_counter()
        RG = (R + G - 1) // G
        specs0, _ = wl.specs(range(seed_base, seed_base + RG))
        lpc, wpb = autotune(specs0, verbose=True, lr_terms=wl.lr_terms)
        log(f"{cfg}: autotune picked lanes_per_chain={lpc} waves_per_block={wpb} in {time.perf_counter() - t0:.1f} s")
    for w in range(warmup):
        seeds = range(1_000_000 + w * R, 1_000_000 + (w + 1) * R)
        t0 = time.perf_counter()
        if slots:
            run_step_queue(wl, seeds, lpc, wpb, slots, graph_iters)
        else:
            bs, _ = make_groups(wl, seeds, lpc, wpb, G)
            run_groups(bs, graph_iters)
            for b in bs:
                b.log_Z()
            del bs
        log(f"{cfg} warmup {w}: {R} runs in {time.perf_counter() - t0:.2f} s")
    torch.cuda.synchronize()
    logZ, truths = [], []
    evals_total = 0
    caps = 0
    t_total = 0.0
    for k in range(steps):
        seeds = range(seed_base + k * R, seed_base + (k + 1) * R)
        if slots:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            z, tr, ev, cp = run_step_queue(wl, seeds, lpc, wpb, slots, graph_iters)
            t1 = time.perf_counter()
            evals_total += ev
            caps += cp
        else:
            bs, tr = make_groups(wl, seeds, lpc, wpb, G)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            run_groups(bs, graph_iters)
            t1 = time.perf_counter()
            z = np.concatenate([b.log_Z() for b in bs])
            evals_total += sum(st.num_likelihood_evaluations for b in bs for st in b.states())
            caps += int(sum(b.cap_hits().sum() for b in bs))
            del bs
        torch.cuda.synchronize()
        t_total += time.perf_counter() - t0
        logZ.append(z)
        truths.append(tr)
        log(f"{cfg} step {k}: {R} runs, sampling {t1 - t0:.3f} s, log Z {time.perf_counter() - t1:.3f} s, total "
            f"{t_total:.2f} s, mean |dlogZ| {np.abs(z[:, 0] - tr).mean():.3f}")
    logZ = np.concatenate(logZ)
    truths = np.concatenate(truths)
    dz = logZ[:, 0] - truths
    return dict(runs_per_s=R * steps / t_total, ms_per_step=1e3 * t_total / steps, t_total=t_total,
                evals_per_s=evals_total / t_total, evals_per_run=evals_total / (R * steps), cap_hits=caps,
                mean_dlogZ=float(dz.mean()), se_mean_dlogZ=float(dz.std(ddof=1) / math.sqrt(dz.size)) if dz.size > 1
                else None, mean_abs_dlogZ=float(np.abs(dz).mean()), frac_abs_dlogZ_lt_0_1=float(np.mean(np.abs(dz) < 0.1)),
                mean_stated_uncert=float(logZ[:, 1].mean()), frac_z_lt_3=float(np.mean(np.abs(dz) / logZ[:, 1] < 3)),
                ensemble_std_logZ=float(logZ[:, 0].std()), lpc=lpc, wpb=wpb, logZ=logZ, truths=truths)


def secondary(cfg, R, steps, graph_iters, slots=0, groups=1, warmup=1):
    """One further configuration measured in the same process (autotuned launch shape, `warmup` untimed
    and `steps` timed steps of R runs in `groups` concurrent groups): its runs/s, evaluations/s, log Z
    statistics and step-cap hits, and the end-to-end algorithmic rate of the step against the roofline
    of its dominant kernel."""
    wl = make_workload(cfg)
    m = measure_chosen(wl, cfg, R, groups, steps, warmup, graph_iters, 4, 4, True, slots=slots)
    out = {k: v for k, v in m.items() if k not in ("logZ", "truths", "t_total")}
    specs, _ = wl.specs(range(1))
    work = m["evals_per_s"] * wl.work_per_eval(specs[0]) if cfg != "c3" else None
    out.update(workload=wl.name, runs_per_step=R, steps=steps, warmup=warmup, groups=groups, dtype=wl.dtype,
               queue_slots=slots)
    if work is not None:
        scale = 1e12 if wl.unit == "TFLOP/s" else 1e9
        out["roofline_end_to_end"] = dict(bound=wl.bound, achieved=work / scale, peak=wl.peak / scale, unit=wl.unit,
                                          frac=work / wl.peak, kernel=wl.kernel,
                                          note="algorithmic work of the whole timed step over its wall time")
    vpe = MEASURED_VALU_PER_EVAL.get((cfg, m["lpc"]))
    if vpe is not None:
        rate = vpe * m["evals_per_s"]
        out["valu_issue_end_to_end"] = dict(
            achieved=rate, peak=PEAK_VALU_ISSUE, unit="wavefront VALU instructions/s", frac=rate / PEAK_VALU_ISSUE,
            valu_per_eval=vpe, note="measured SQ_INSTS_VALU per evaluation times the step's evaluations per second")
    return out


def secondary_in_child(key, cfg, R, steps, warmup, groups, graph_iters):
    """secondary() in a child process (python bench.py --emit-secondary): the line's JSON result."""
    import subprocess
    cmd = [sys.executable, os.path.abspath(__file__), "--config", cfg, "--runs", str(R), "--steps", str(steps),
           "--warmup", str(warmup), "--groups", str(groups), "--graph-iters", str(graph_iters),
           "--emit-secondary", key]
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=None, text=True, timeout=600)
    for ln in p.stdout.splitlines():
        if not ln.startswith('{"emit_secondary"'):
            print(ln, file=sys.stderr, flush=True)  # the child's autotuner lines
    if p.returncode != 0:
        raise RuntimeError(f"secondary {key}: child exited with {p.returncode}")
    lines = [ln for ln in p.stdout.splitlines() if ln.startswith('{"emit_secondary"')]
    return json.loads(lines[-1])["result"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--runs", type=int, default=None,
                    help="runs per step (chosen form); default c2: 1024, c1: 1024, c3: 2048, c4: 256, c5: 8")
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--shells", type=int, default=3, help="c5: shells per partial run")
    ap.add_argument("--full-runs", type=int, default=0, help="c5: also run this many whole runs to termination")
    ap.add_argument("--graph-iters", type=int, default=8)
    ap.add_argument("--groups", type=int, default=1,
                    help="chosen form: the runs of a step in this many concurrent groups, each on its own stream "
                         "(the roofline's instrumented launches then cover one group)")
    ap.add_argument("--slots", type=int, default=-1,
                    help="stream the runs of a step through this many device slots (engine.run_queue); "
                         "0 = one batch (the default)")
    ap.add_argument("--lpc", type=int, default=4)
    ap.add_argument("--wpb", type=int, default=4)
    ap.add_argument("--lr-terms", type=int, default=3, choices=[3, 4])
    ap.add_argument("--plain-runs", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--emit-secondary", default=None,
                    help="print one secondary line of --config (runs, steps, warmup, groups as given) as JSON")
    ap.add_argument("--no-secondary", action="store_true",
                    help="skip the one-step measurements of the other configurations")
    ap.add_argument("--no-autotune", action="store_true", help="use --lpc/--wpb instead of the autotuner")
    ap.add_argument("--profile-pass", action="store_true",
                    help="only the instrumented shell-by-shell pass (for rocprofv3 kernel statistics)")
    ap.add_argument("--to-results-scale", type=int, default=0,
                    help="time the evidence scan, count_crossed_edges and the stable argsort at this many samples")
    ap.add_argument("--copy-calib", type=int, default=0,
                    help="with --profile-pass / --to-results-scale: also copy a buffer of this many MiB first "
                         "(counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU (--gpus 1): the runs of a step are independent, so several GPUs "
              "would each take a share of them in a process of their own, and that form is not measured",
              file=sys.stderr)
        sys.exit(2)
    import torch
    torch.cuda.set_device(0)
    if args.copy_calib and (args.profile_pass or args.to_results_scale):
        n = args.copy_calib * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b
        log(f"copy calibration: 3 x {args.copy_calib} MiB read + written")
    if args.to_results_scale:
        print(json.dumps(to_results_scale(args.to_results_scale, args.steps)), flush=True)
        return
    if args.emit_secondary:
        res = secondary(args.config, args.runs, args.steps, args.graph_iters, slots=QUEUE_SLOTS.get(args.config, 0),
                        groups=max(1, args.groups), warmup=args.warmup)
        print(json.dumps({"emit_secondary": args.emit_secondary, "result": res}), flush=True)
        return
    wl = make_workload(args.config)
    if args.config == "c4":
        wl.lr_terms = args.lr_terms
    # runs per step: c2 1024 (512: 920 runs/s, 768: 939, 1024: 952 on one box; the shell kernel's tail and
    # the per-run kernels amortise over more runs); c3 1024 (256: 54.7, 512: 65.9, 1024: 74.9)
    R = args.runs or {"c1": 1024, "c2": 1024, "c3": 2048, "c4": 256, "c5": 8}.get(args.config, 128)
    G = max(1, args.groups)
    RG = (R + G - 1) // G  # runs per group: the batch one launch of the dominant kernel covers
    seed_base = 0
    if args.config == "c5":
        t0 = time.perf_counter()
        res = precision_at_scale(wl, R, args.shells, 4, args.wpb, args.steps)
        wall = time.perf_counter() - t0
        full = full_runs_at_scale(wl, args.full_runs, 1, args.graph_iters) if args.full_runs else None
        bf, fp = res["bf16x3"], res["fp32"]
        roof = dict(bound="mfma", achieved=bf["achieved_TFLOPs"], peak=PEAK_BF16_MFMA / 1e12, unit="TFLOP/s",
                    frac=bf["achieved_TFLOPs"] * 1e12 / PEAK_BF16_MFMA, traffic=None, kernel=wl.kernel,
                    avg_launch_ms=bf["avg_launch_ms"], evals_per_launch=bf["evals_per_launch"])
        print(json.dumps({
            "metric": "likelihood evals/s of partial runs (BF16x3 vs plain FP32), 1 GPU", "value": bf["evals_per_s"],
            "unit": "evals/s", "n_gpus": 1, "steps": args.shells, "warmup": 0, "ms_per_step": bf["avg_launch_ms"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "bf16x3 likelihood GEMM (MFMA, f32 accumulate)", "data": "synthetic (seeded numpy)",
            "config": {"workload": wl.name, "runs_per_step": R, "shells": args.shells, "lanes_per_chain": 4,
                       "waves_per_block": args.wpb},
            "bf16x3": bf, "plain_fp32": fp, "bf16x3_vs_fp32": bf["evals_per_s"] / fp["evals_per_s"],
            "log_Z_true": float(wl.log_Z), "wall_s": wall, "roofline": roof, "cpu_baseline": None,
            "full_runs": full}), flush=True)
        return
    if args.profile_pass:
        durs, work, evs = instrumented_slice_timing(wl, R, args.lpc, args.wpb, seed0=90000)
        print(json.dumps({"profile_pass": True, "config": args.config, "runs": R, "lanes_per_chain": args.lpc,
                          "waves_per_block": args.wpb, "sample_launches": int(durs.size),
                          "avg_launch_ms": float(durs.mean() * 1e3), "evals_per_launch": float(evs.mean()),
                          "achieved": float(work.mean() / durs.mean()), "unit": wl.unit[:-2] + "/s",
                          "lane_utilisation": instrumented_slice_timing.lane_util}), flush=True)
        return

    # ---------------- chosen form: R runs per step
    slots = QUEUE_SLOTS.get(args.config, 0) if args.slots < 0 else args.slots
    m = measure_chosen(wl, args.config, R, G, args.steps, args.warmup, args.graph_iters, args.lpc, args.wpb,
                       not args.no_autotune, seed_base, slots=slots)
    args.lpc, args.wpb = m["lpc"], m["wpb"]
    ms_per_step = m["ms_per_step"]
    runs_per_s = m["runs_per_s"]
    dz = m["logZ"][:, 0] - m["truths"]

    # ---------------- plain form: one run at a time, host sync per shell, FP32 likelihood GEMM
    tp = 0.0
    pz = []
    for k in range(args.plain_runs):
        b, tr = make_batch(wl, [seed_base + k], args.lpc, args.wpb, plain=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pz.append(run_batch(b, 0)[0][0] - tr[0])
        torch.cuda.synchronize()
        tp += time.perf_counter() - t0
        log(f"plain run {k}: {tp:.2f} s so far")
        del b
    plain_runs_per_s = args.plain_runs / tp if args.plain_runs else None

    # ---------------- roofline of the dominant kernel (hipns_shell_sample): the launches of one group of
    # the timed step (R runs per launch with the default single group), shell by shell with HIP events
    durs, work, evs = instrumented_slice_timing(wl, RG, args.lpc, args.wpb, seed0=90000)
    log(f"instrumented: {durs.size} sample launches, mean {durs.mean() * 1e3:.3f} ms")
    avg_dur = float(durs.mean())
    achieved = float(work.mean() / avg_dur)
    scale = 1e12 if wl.unit == "TFLOP/s" else 1e9
    launches_per_step = durs.size * G
    roofline = dict(bound=wl.bound, achieved=achieved / scale, peak=wl.peak / scale, unit=wl.unit,
                    frac=achieved / wl.peak, traffic=MEASURED_TRAFFIC.get((args.config, RG)),
                    traffic_unit="bytes per launch", kernel=wl.kernel, avg_launch_ms=avg_dur * 1e3,
                    launches=int(durs.size), evals_per_launch=float(evs.mean()),
                    evals_per_s_in_kernel=float(evs.sum() / durs.sum()),
                    lane_utilisation=instrumented_slice_timing.lane_util,
                    kernel_ms_per_step=avg_dur * 1e3 * launches_per_step,
                    end_to_end_achieved=m["evals_per_s"] * wl.work_per_eval(wl.specs(range(1))[0][0]) / scale)
    valu = MEASURED_VALU.get((args.config, RG))
    if valu is not None:
        # the kernel's VALU issue: SQ_INSTS_VALU per launch (rocprofv3 --pmc on the same launch shape) over the
        # live average launch duration, against one wavefront instruction per 4 cycles per SIMD
        rate = valu / avg_dur
        roofline["valu_ceiling"] = dict(insts_per_launch=valu, achieved=rate, peak=PEAK_VALU_ISSUE,
                                        unit="wavefront VALU instructions/s", frac=rate / PEAK_VALU_ISSUE)

    cpu = None
    if not args.no_cpu:
        specs1, _ = wl.specs([seed_base])
        cpu = cpu_baseline(wl, specs1[0], m["evals_per_run"])
    log("cpu baseline", cpu)
    sec = {}
    if not args.no_secondary and args.config == "c2":
        # (line, workload, runs per step, timed steps, warm-up steps, concurrent groups): the mixture
        # batch at the 256 runs of its configuration, in four concurrent groups of 64 (one per hardware
        # queue: 1 group 153 runs/s, 2 168, 4 175, 8 109 on one box), and at 2048 in one (219 runs/s; 4
        # groups 215), whose 9 s step runs cold: first-launch costs are below 1 % of it
        # Each line runs in a child process of its own, as the workload would run alone: in this process,
        # after the headline's batches, the mixture batch's second step ran 38 % slower (1.85 s against
        # 1.33 s for the same runs in a fresh process)
        for key, cfg, Rs, st, wu, gr in (("c1", "c1", 1024, 8, 1, 1), ("c4", "c4", 256, 1, 1, 2),
                                         ("c3", "c3", 256, 2, 1, 4), ("c3_2048", "c3", 2048, 1, 0, 1)):
            t0 = time.perf_counter()
            # a line that fails is recorded as such; the headline line is printed in any case
            try:
                sec[key] = secondary_in_child(key, cfg, Rs, st, wu, gr, args.graph_iters)
                log(f"secondary {key}: {sec[key]['runs_per_s']:.2f} runs/s in {time.perf_counter() - t0:.1f} s")
            except Exception as e:  # noqa: BLE001
                sec[key] = dict(error=f"{type(e).__name__}: {e}")
                log(f"secondary {key} failed: {e}")
        # precision at scale (65 536 x 128 regression): three shells of 8 partial runs in each product form
        # (a whole run needs ~2 100 shells), their evaluation rate and the accepted points' logL error
        # against the FP64 likelihood
        t0 = time.perf_counter()
        try:
            wl5 = make_workload("c5")
            res5 = precision_at_scale(wl5, 8, 3, 4, 4, 1)
            bf, fp = res5["bf16x3"], res5["fp32"]
            sec["c5"] = dict(workload=wl5.name, runs_per_step=8, shells=3, evals_per_s=bf["evals_per_s"],
                             plain_fp32_evals_per_s=fp["evals_per_s"],
                             bf16x3_vs_fp32=bf["evals_per_s"] / fp["evals_per_s"],
                             logL_rel_err_max=bf["logL_rel_err_max"],
                             unrefined_logL_rel_err_max=bf.get("unrefined_logL_rel_err_max"),
                             plain_fp32_logL_rel_err_max=fp["logL_rel_err_max"],
                             roofline=dict(bound="mfma", achieved=bf["achieved_TFLOPs"], peak=PEAK_BF16_MFMA / 1e12,
                                           unit="TFLOP/s", frac=bf["achieved_TFLOPs"] * 1e12 / PEAK_BF16_MFMA,
                                           kernel=wl5.kernel, avg_launch_ms=bf["avg_launch_ms"]))
            log(f"secondary c5: {bf['evals_per_s'] / 1e6:.1f} M evals/s in {time.perf_counter() - t0:.1f} s")
        except Exception as e:  # noqa: BLE001
            sec["c5"] = dict(error=f"{type(e).__name__}: {e}")
            log(f"secondary c5 failed: {e}")
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
        "dtype": wl.dtype,
        "data": "synthetic (seeded numpy)",
        "config": {"workload": wl.name, "runs_per_step": R, "groups": G, "queue_slots": slots,
                   "graph_iters": args.graph_iters,
                   "lanes_per_chain": args.lpc, "waves_per_block": args.wpb},
        "likelihood_evals_per_s": m["evals_per_s"],
        "evals_per_run": m["evals_per_run"],
        "cap_hits": m["cap_hits"],
        "plain_form_runs_per_s": plain_runs_per_s,
        "chosen_vs_plain": runs_per_s / plain_runs_per_s if plain_runs_per_s else None,
        "mean_dlogZ": float(dz.mean()),
        "se_mean_dlogZ": m["se_mean_dlogZ"],
        "mean_abs_dlogZ": float(np.abs(dz).mean()),
        "frac_abs_dlogZ_lt_0.1": float(np.mean(np.abs(dz) < 0.1)),
        "mean_stated_uncert": m["mean_stated_uncert"],
        "frac_z_lt_3": m["frac_z_lt_3"],
        "plain_dlogZ": [float(z) for z in pz],
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    oracle_ref = ORACLE_ENSEMBLE_KEY.get(args.config)
    if oracle_ref:
        out["oracle_ensemble"] = oracle_ensemble(oracle_ref)
    if args.config in ("c1", "c2", "c4"):
        out["log_Z_true"] = float(m["truths"][0])
        out["ensemble_std_logZ"] = m["ensemble_std_logZ"]
    if sec:
        out["secondary"] = sec
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
