"""Writes tests/golden/oracle_logZ_ensembles.json: log Z of many runs of the CPU oracle (the FP64
restatement of the reference algorithm) on the benchmark workloads, against the analytic value.

The fixture is the like-for-like partner of the device ensembles: a GPU test runs the same
workload on the device and checks that both mean Delta log Z agree within their combined standard
error.  Workloads:
  mvn_d10   MVN D=10, num_live_points=300, s=5 (400 runs)
  mvn_d32   MVN D=32, num_live_points=2048, s=4 (400 runs)
  linreg    regression 4096 x 16, sigma=0.3, tau=1, num_live_points=1024, s=5 (400 runs); the
            likelihood is evaluated in its Gaussian-in-w form (same function, O(D^2) per call)
  gmm       the mixture batch's first 256 run specifications (seeds 0..255: D, num_live_points,
            means and widths drawn from the seed), one oracle run per specification
  linreg_s10, linreg_s20   the regression workload with s = 10 and 20 slices per dimension (200 runs
            each): how the algorithm's log Z bias at s = 5 changes with longer chains
  gmm_dir16 the mixture runs of gmm with the slice directions drawn as the device draws them (16-bit
            uniform pairs), to compare the two direction forms
Each oracle run uses its own stream (seed 7_000_000 + index); the device uses Philox streams, so
the runs are statistically, not bitwise, matched.

  python tests/golden/make_oracle_ensembles.py [workload ...] [--procs 8] [--extend K]

--extend K adds K runs (the next K indices) to a workload already in the file and recomputes its
summary over all of its runs.
"""
import argparse
import json
import math
import os
import sys
import time
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
OUT = os.path.join(HERE, "oracle_logZ_ensembles.json")

WORKLOADS = {
    "mvn_d10": dict(runs=400),
    "mvn_d32": dict(runs=400),
    "linreg": dict(runs=400),
    "gmm": dict(runs=256),
    "linreg_s10": dict(runs=200),
    "linreg_s20": dict(runs=200),
    "gmm_dir16": dict(runs=1024),
}


def _setup(name, i):
    """(oracle spec, N, num_slices, max_samples, log Z analytic, logL constant, D) of run i."""
    import problems as P
    if name in ("mvn_d10", "mvn_d32"):
        D, N, s = (10, 300, 5) if name == "mvn_d10" else (32, 2048, 4)
        _, log_Z, mu, Sig = P.mvn_problem(D, seed=0)
        return P.mvn_oracle_spec(D, mu, Sig), N, s * D, N * 120, log_Z, 0.0, D
    if name.startswith("linreg"):
        s = 5 if name == "linreg" else int(name.split("_s")[1])
        _, log_Z, lik = P.linreg_problem(4096, 16, seed=0, sigma=0.3, tau=1.0)
        spec, const = P.linreg_mvn_oracle_spec(lik, 1.0)
        return spec, 1024, 16 * s, 1024 * 200, log_Z, const, 16
    if name in ("gmm", "gmm_dir16"):
        rng = np.random.default_rng(i)
        D = int(rng.integers(4, 65))
        N = int(rng.integers(128, 1025))
        _, log_Z, means, sig = P.gmm_problem(D, seed=10_000 + i)
        return P.gmm_oracle_spec(D, means, sig), N, 5 * D, N * 300, log_Z, 0.0, D
    raise KeyError(name)


def _one(args):
    import oracle as O
    name, i = args
    spec, N, ns, ms, log_Z, const, D = _setup(name, i)
    t0 = time.perf_counter()
    r = O.ns_run(spec, num_live_points=N, num_slices=ns, max_samples=ms, seed=7_000_000 + i, num_threads=1,
                 dir16=name.endswith("_dir16"))
    return dict(i=i, D=D, N=N, dlogZ=r["log_Z_mean"] + const - log_Z, uncert=r["log_Z_uncert"],
                evals=int(r["total_likelihood_evaluations"]), reason=int(r["termination_reason"]),
                wall_s=time.perf_counter() - t0)


def summarise(rows):
    dz = np.array([r["dlogZ"] for r in rows])
    un = np.array([r["uncert"] for r in rows])
    n = dz.size
    return dict(runs=n, mean_dlogZ=float(dz.mean()), se_mean_dlogZ=float(dz.std(ddof=1) / math.sqrt(n)),
                std_dlogZ=float(dz.std(ddof=1)), mean_abs_dlogZ=float(np.abs(dz).mean()),
                mean_stated_uncert=float(un.mean()), frac_abs_dlogZ_lt_0_1=float(np.mean(np.abs(dz) < 0.1)),
                frac_z_lt_3=float(np.mean(np.abs(dz) / un < 3)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("workloads", nargs="*", default=list(WORKLOADS))
    ap.add_argument("--procs", type=int, default=8)
    ap.add_argument("--extend", type=int, default=0)
    a = ap.parse_args()
    data = json.load(open(OUT)) if os.path.exists(OUT) else {}
    for name in a.workloads:
        old = data.get(name, {}).get("runs", []) if a.extend else []
        first = max((r["i"] for r in old), default=-1) + 1
        n = a.extend if a.extend else WORKLOADS[name]["runs"]
        t0 = time.perf_counter()
        rows = []
        with Pool(a.procs) as pool:
            for r in pool.imap_unordered(_one, [(name, i) for i in range(first, first + n)]):
                rows.append(r)
                if len(rows) % 20 == 0:
                    print(f"{name}: {len(rows)}/{n} runs, {time.perf_counter() - t0:.0f} s", flush=True)
        rows = list(old) + rows
        rows.sort(key=lambda r: r["i"])
        s = summarise(rows)
        data[name] = dict(summary=s, runs=[dict(i=r["i"], D=r["D"], N=r["N"], dlogZ=round(r["dlogZ"], 6),
                                                uncert=round(r["uncert"], 6), evals=r["evals"]) for r in rows])
        print(name, json.dumps(s), flush=True)
        with open(OUT, "w") as f:
            json.dump(data, f, indent=1)


if __name__ == "__main__":
    main()
