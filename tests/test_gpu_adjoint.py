"""Evidence maximisation for Bayesian linear regression: the hand-written adjoint of the batched
likelihood (hipns_linreg_stats behind a torch.autograd.Function) against finite differences of the
same objective and against the analytic linear-Gaussian log Z and gradient."""
import math

import numpy as np
import pytest

import jaxns_amd as J
from jaxns_amd.evidence_maximisation import (EvidenceMaximisation, LinearRegressionEvidenceModel,
                                             LinRegAdjointStats, log_evidence)

from problems import linreg_problem

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _e_step(fam, theta, seed=3, nlive=512):
    ns = J.NestedSampler(fam(theta), num_live_points=nlive, s=4, max_samples=nlive * 300)
    reason, state = ns(seed)
    return ns.to_results(reason, state)


def test_adjoint_vs_finite_differences_and_analytic():
    import torch
    say("adjoint: E-step")
    _, _, lik = linreg_problem(1024, 8, seed=11, sigma=0.3)
    fam = LinearRegressionEvidenceModel(lik.data)
    theta0 = np.array([math.log(0.3), math.log(1.0)])
    res = _e_step(fam, theta0)
    say("E-step log Z", res.log_Z_mean, "+-", res.log_Z_uncert, "analytic", fam.log_Z_analytic(theta0),
        "samples", res.U_samples.shape)
    lw = res.log_dp_mean - res.log_L_samples + res.log_Z_mean
    stats = LinRegAdjointStats(fam, res.U_samples)
    lw_t = torch.from_numpy(np.asarray(lw, dtype=np.float64)).cuda()
    # forward at theta0 reproduces the sampler's likelihood values and log Z
    logL, rss, yr = stats(theta0)
    rel = np.abs(logL.cpu().numpy() - res.log_L_samples) / np.maximum(np.abs(res.log_L_samples), 1.0)
    say("stats logL vs sampler logL: max rel", rel.max())
    assert rel.max() < 1e-9
    th = torch.tensor(theta0, dtype=torch.float64, requires_grad=True, device="cuda")
    lz = log_evidence(th, stats, lw_t)
    lz.backward()
    g = th.grad.cpu().numpy()
    assert abs(float(lz.detach()) - res.log_Z_mean) < 1e-6
    # adjoint vs central differences of the objective evaluated in FP64 on the host at the same
    # samples and weights.  (Differences of the device objective itself are not a fair check: its
    # 1e-4 nat approximation error moves with theta, which biases d/d log tau by ~0.5 %.)
    from scipy.special import logsumexp, ndtri
    Uh = np.asarray(res.U_samples, dtype=np.float64)
    lw_h = np.asarray(lw, dtype=np.float64)

    def log_Z_host(th):
        sg, tau = np.exp(th)
        R = lik.y[None, :] - (tau * ndtri(Uh)) @ lik.X.T
        L = -0.5 * lik.N * np.log(2 * np.pi * sg * sg) - 0.5 * np.sum(R * R, axis=1) / sg ** 2
        return logsumexp(L + lw_h)

    h = 1e-5
    fd = []
    for k in range(2):
        e = np.zeros(2)
        e[k] = h
        fd.append((log_Z_host(theta0 + e) - log_Z_host(theta0 - e)) / (2 * h))
    say("adjoint", g, "finite differences (FP64 host)", fd)
    np.testing.assert_allclose(g, fd, rtol=2e-3, atol=1e-3)
    # vs the analytic gradient, within the weighted-sample error of the posterior mean
    ga = fam.grad_log_Z_analytic(theta0)
    p = np.exp(res.log_dp_mean - np.max(res.log_dp_mean))
    p /= p.sum()
    s2 = 0.09
    dls = -1024 + rss.cpu().numpy() / s2
    dlt = (yr.cpu().numpy() - rss.cpu().numpy()) / s2
    sd = np.array([np.sqrt(p @ (dls - p @ dls) ** 2), np.sqrt(p @ (dlt - p @ dlt) ** 2)])
    tol = 5 * sd / math.sqrt(res.ESS) + 0.2
    say("analytic", ga, "tolerance", tol)
    assert np.all(np.abs(g - ga) < tol)


def test_evidence_maximisation_improves_analytic_evidence():
    """EM on (log sigma, log tau): sigma reaches the analytic optimum given the final tau, tau moves
    toward its optimum and the analytic evidence rises.  (With the samples fixed in the unit cube,
    w = tau q(U) moves every sample when tau changes, so each M-step moves tau only part way; the
    reference's objective has the same property.)"""
    from scipy.optimize import minimize_scalar
    say("evidence maximisation")
    _, _, lik = linreg_problem(512, 4, seed=12, sigma=0.5)
    theta0 = np.array([math.log(1.0), math.log(3.0)])
    fam = LinearRegressionEvidenceModel(lik.data, params=theta0)
    em = EvidenceMaximisation(fam, ns_kwargs=dict(num_live_points=512, s=4, max_samples=512 * 300), gtol=1e-3,
                              log_Z_ftol=0.0, log_Z_atol=0.0)
    res, params = em.train(num_steps=4)
    lt_opt = minimize_scalar(lambda lt: -fam.log_Z_analytic((params[0], lt)), bounds=(-5, 5), method="bounded").x
    ls_opt = minimize_scalar(lambda ls: -fam.log_Z_analytic((ls, params[1])), bounds=(-5, 5), method="bounded").x
    z0, z1 = fam.log_Z_analytic(theta0), fam.log_Z_analytic(params)
    say("EM params", params, "sigma optimum given tau", ls_opt, "tau optimum given sigma", lt_opt,
        "analytic log Z", z0, "->", z1)
    assert abs(params[0] - ls_opt) < 0.02
    assert abs(params[1] - lt_opt) < abs(theta0[1] - lt_opt)
    assert z1 > z0 + 1.0
