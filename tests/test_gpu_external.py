"""Likelihoods given as torch callables (the general form): one-step sampler launches with the
likelihood evaluated between them (hipns_ext_step), against the fused kernels and analytic evidence."""
import math

import numpy as np
import pytest

import jaxns_amd as J
from jaxns_amd import distributions as tfpd
from jaxns_amd.engine import DeviceBatch, RunSpec

from problems import mvn_problem

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _gauss_1d():
    def prior_model():
        x = yield J.Prior(tfpd.Uniform(low=0.0, high=1.0), name="x")
        return x

    lik = J.GaussianMixtureLikelihood(np.array([[0.3]]), np.array([0.05]))
    # the same arithmetic as the fused mixture kernel: FP32 residual and square, FP64 assembly
    lc = float(np.float32(lik.logc[0]))
    i2 = float(np.float32(0.5 / 0.05 ** 2))
    m = float(np.float32(0.3))

    def fn(x):
        r = x.float() - m
        return lc - (r * r).double() * i2

    return J.Model(prior_model, lik), J.Model(prior_model, fn)


@pytest.mark.parametrize("k", [0, 1])
def test_external_matches_fused_bitwise(k):
    """A 1-D likelihood whose device arithmetic the callable reproduces exactly: the stepwise path and
    the fused kernel give bit-identical samples, evaluation counts and log Z (initial live points,
    slice shells with shrinks, phantoms, termination and results)."""
    fused_model, ext_model = _gauss_1d()
    out = []
    for model in (fused_model, ext_model):
        specs = [RunSpec(model=model, num_live_points=64, num_slices=5, k_phantom=k, max_samples=64 * 80,
                         seed=s) for s in (3, 4)]
        b = DeviceBatch(specs, lanes_per_chain=2, waves_per_block=2)
        b.run(dict(dlogZ=math.log(1.0 + 1e-3), max_samples=1e18), graph_iters=0)
        out.append((b, b.log_Z()))
    (bf, zf), (be, ze) = out
    say("log Z fused", zf[:, 0], "stepwise", ze[:, 0])
    for r in range(2):
        rf, re_ = bf.to_results(r), be.to_results(r)
        assert rf.total_num_samples == re_.total_num_samples
        assert np.array_equal(rf.log_L_samples, re_.log_L_samples)
        assert np.array_equal(rf.U_samples, re_.U_samples)
        assert rf.total_num_likelihood_evaluations == re_.total_num_likelihood_evaluations
    assert np.array_equal(zf, ze)


def test_external_mvn_logZ():
    """A callable multivariate-normal likelihood (FP64 torch) through NestedSampler: log Z against
    the analytic value."""
    import torch
    model, log_Z, mu, Sig = mvn_problem(6, seed=4)
    L = torch.linalg.cholesky(torch.tensor(Sig, dtype=torch.float64, device="cuda"))
    mu_t = torch.tensor(mu, dtype=torch.float64, device="cuda")
    c = -0.5 * 6 * math.log(2 * math.pi) - float(torch.log(torch.diagonal(L)).sum())

    def log_likelihood(x):
        z = torch.linalg.solve_triangular(L, (x - mu_t).T, upper=False)
        return c - 0.5 * (z * z).sum(0)

    ext = J.Model(model.prior_model, log_likelihood)
    ns = J.NestedSampler(ext, num_live_points=200, s=4, max_samples=200 * 150)
    reason, state = ns(7)
    res = ns.to_results(reason, state)
    say("log Z", res.log_Z_mean, "+-", res.log_Z_uncert, "analytic", log_Z, "evals",
        res.total_num_likelihood_evaluations)
    assert abs(res.log_Z_mean - log_Z) < 4 * res.log_Z_uncert + 0.1
    assert reason & 4


def test_external_two_priors_arguments():
    """The callable receives the prior values the generator returns, one argument per prior."""
    import torch

    def prior_model():
        a = yield J.Prior(tfpd.Normal(loc=0.0, scale=1.0), name="a")
        b = yield J.Prior(tfpd.Uniform(low=-2.0 * np.ones(2), high=2.0 * np.ones(2)), name="b")
        return a, b

    seen = {}

    def log_likelihood(a, b):
        seen["shapes"] = (tuple(a.shape), tuple(b.shape))
        return -0.5 * (a ** 2 + (b ** 2).sum(-1))

    m = J.Model(prior_model, log_likelihood)
    specs = [RunSpec(model=m, num_live_points=50, num_slices=6, max_samples=50 * 60, seed=1)]
    b = DeviceBatch(specs, lanes_per_chain=2, waves_per_block=1)
    b.run(dict(dlogZ=math.log(1.0 + 1e-2), max_samples=1e18))
    B = seen["shapes"][0][0]
    assert seen["shapes"] == ((B,), (B, 2))
    z = b.log_Z()[0]
    # analytic: E_prior[exp(-a^2/2)] * E_uniform[exp(-|b|^2/2)]
    from scipy.special import erf
    za = 1.0 / math.sqrt(2.0)
    zb = (math.sqrt(math.pi / 2) * erf(2 / math.sqrt(2)) / 2.0) ** 2
    say("log Z", z, "analytic", math.log(za * zb))
    assert abs(z[0] - math.log(za * zb)) < 4 * z[1] + 0.1


def test_gradient_guided_mvn_logZ():
    """gradient_guided=True (uni_slice_sampler.py:255-271) in the stepwise form, the gradient from torch
    autograd of a callable likelihood (the MVN written in torch): log Z against the analytic value, and
    every slice-sampled point spent at least two evaluations per slice (its accepted proposal plus the
    gradient at it)."""
    import torch
    fused, log_Z, mu, Sig = mvn_problem(5, seed=6)
    P = torch.linalg.inv(torch.from_numpy(Sig))
    mu_t = torch.from_numpy(mu)
    c = -0.5 * 5 * math.log(2 * math.pi) - 0.5 * float(np.linalg.slogdet(Sig)[1])

    def log_likelihood(x):
        r = x - mu_t.to(x.device)
        return c - 0.5 * ((r @ P.to(x.device)) * r).sum(-1)

    model = J.Model(fused.prior_model, log_likelihood)
    ns = J.NestedSampler(model, num_live_points=150, s=3, max_samples=150 * 150, gradient_guided=True)
    reason, state = ns(9)
    res = ns.to_results(reason, state)
    ev = np.asarray(res.num_likelihood_evaluations_per_sample)
    say("log Z", res.log_Z_mean, "+-", res.log_Z_uncert, "analytic", log_Z, "min evals (slice samples)",
        ev[ev > 1].min())
    assert abs(res.log_Z_mean - log_Z) < 4 * res.log_Z_uncert + 0.15
    # initial points took >= 1 evaluation; slice-sampled ones 2 per slice at least
    assert (ev[ev > 1] >= 2 * ns.num_slices).all()
