"""Ensemble log Z of the device against the oracle, like for like.

tests/golden/oracle_logZ_ensembles.json holds log Z - analytic log Z of many oracle runs (the FP64
restatement of the reference algorithm, tests/golden/make_oracle_ensembles.py) on the benchmark
workloads.  Here the device runs seeded ensembles of the same workloads, and the two mean Delta log Z
must agree within 3 combined standard errors: a bias of the device that the reference algorithm does
not have fails, while a bias that the algorithm itself has (and the oracle shows) passes.
"""
import json
import math
import os

import numpy as np
import pytest

from problems import gmm_problem, linreg_problem, mvn_problem

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "oracle_logZ_ensembles.json")


def say(*a):
    print(*a, flush=True)


def _fixture(name):
    if not os.path.exists(FIXTURE):
        pytest.skip("no oracle ensemble fixture")
    data = json.load(open(FIXTURE))
    if name not in data:
        pytest.skip(f"no oracle ensemble for {name}")
    return data[name]


def _device_ensemble(name, R, seed0, groups=1):
    from jaxns_amd.engine import DeviceBatch, RunSpec, run_concurrent
    specs, truths = [], []
    if name in ("mvn_d10", "mvn_d32"):
        D, N, s = (10, 300, 5) if name == "mvn_d10" else (32, 2048, 4)
        model, log_Z, _, _ = mvn_problem(D, seed=0)
        specs = [RunSpec(model=model, num_live_points=N, num_slices=s * D, max_samples=N * 120, seed=seed0 + i)
                 for i in range(R)]
        truths = [log_Z] * R
        lpc = 4
    elif name.startswith("linreg"):
        # s = 5 slices per dimension (the benchmark workload), or 10 and 20 (the oracle's s-sweep)
        s = {"linreg": 5, "linreg_s10": 10, "linreg_s20": 20}[name]
        model, log_Z, _ = linreg_problem(4096, 16, seed=0, sigma=0.3, tau=1.0)
        specs = [RunSpec(model=model, num_live_points=1024, num_slices=16 * s, max_samples=1024 * 200,
                         seed=seed0 + i) for i in range(R)]
        truths = [log_Z] * R
        lpc = 2
    else:
        for i in range(R):
            rng = np.random.default_rng(i)
            D = int(rng.integers(4, 65))
            N = int(rng.integers(128, 1025))
            model, log_Z, _, _ = gmm_problem(D, seed=10_000 + i)
            specs.append(RunSpec(model=model, num_live_points=N, num_slices=5 * D, max_samples=N * 300,
                                 seed=seed0 + i))
            truths.append(log_Z)
        lpc = 8  # the bench's launch shape for the mixture batch (row stride 64 covers the 8-lane instance)
    cut = np.linspace(0, R, groups + 1).round().astype(int)
    batches = [DeviceBatch(specs[cut[g]:cut[g + 1]], lanes_per_chain=lpc, waves_per_block=2) for g in range(groups)]
    terms = [dict(dlogZ=math.log(1.0 + 1e-3), max_samples=float(max(s.max_samples for s in b.specs)))
             for b in batches]
    if groups == 1:
        batches[0].run(terms[0], graph_iters=8)
    else:
        run_concurrent(batches, terms, graph_iters=8)
    z = np.concatenate([b.log_Z() for b in batches])
    caps = int(sum(b.cap_hits().sum() for b in batches))
    return z[:, 0] - np.array(truths), z[:, 1], caps


@pytest.mark.parametrize("name,R", [("mvn_d10", 400), ("mvn_d32", 512), ("linreg", 256), ("linreg_s10", 256),
                                    ("linreg_s20", 256), ("gmm", 256)])
def test_device_ensemble_matches_oracle(name, R):
    fx = _fixture(name)["summary"]
    say(f"ensemble {name}: {R} device runs against {fx['runs']} oracle runs")
    dz, unc, caps = _device_ensemble(name, R, seed0=500_000, groups=2 if R >= 256 else 1)
    m, se = float(dz.mean()), float(dz.std(ddof=1) / math.sqrt(dz.size))
    say(f"  device mean dlogZ {m:+.4f} +- {se:.4f} (std {dz.std():.4f}, stated {unc.mean():.4f}); "
        f"oracle {fx['mean_dlogZ']:+.4f} +- {fx['se_mean_dlogZ']:.4f} (std {fx['std_dlogZ']:.4f}, stated "
        f"{fx['mean_stated_uncert']:.4f}); step-cap hits {caps}")
    assert caps == 0
    tol = 3.0 * math.sqrt(se ** 2 + fx["se_mean_dlogZ"] ** 2)
    assert abs(m - fx["mean_dlogZ"]) < tol, (m, fx["mean_dlogZ"], tol)
    # the spread of the runs is the algorithm's too: the ratio of the two standard deviations is
    # within the sampling error of two ensembles of a few hundred runs
    r = dz.std(ddof=1) / fx["std_dlogZ"]
    assert 0.75 < r < 1.33, r
