"""The reference's end-to-end fixtures (src/jaxns/tests/conftest.py: basic, basic_with_obj, basic2,
basic3, plateau, basic_mvn) with their NestedSampler settings, checked as in
src/jaxns/tests/test_nested_sampler.py:9-37: the sample_evidence ensemble mean within 3 sigma of
the true log Z, log_Z_mean within 3 sigma of the ensemble mean, and log_Z_uncert consistent with
the ensemble spread.  Likelihoods are torch callables (the general path); the true values are
closed forms or quadrature."""
import math

import numpy as np
import pytest
from scipy import integrate
from scipy.special import erf
from scipy.stats import multivariate_normal

import jaxns_amd as J
from jaxns_amd import distributions as tfpd

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _basic():
    def prior_model():
        x = yield J.Prior(tfpd.Uniform(low=0, high=1), name="x")
        return x

    return J.Model(prior_model, lambda x: -x ** 2), math.log(math.sqrt(math.pi) / 2 * erf(1.0)), 1000


def _basic_with_obj():
    class Obj:
        def __init__(self, x):
            self.x = x

    def prior_model():
        x = yield J.Prior(tfpd.Uniform(low=0, high=1), name="x")
        return Obj(x)

    return J.Model(prior_model, lambda obj: -obj.x ** 2), math.log(math.sqrt(math.pi) / 2 * erf(1.0)), 1000


def _basic2():
    import torch

    def prior_model():
        x = yield J.Prior(tfpd.Uniform(low=0, high=1))
        return x

    return J.Model(prior_model, lambda x: torch.log(1.0 - x ** 2)), math.log(1.0 - 1.0 / 3.0), 1000


def _basic3():
    def prior_model():
        x = yield J.Prior(tfpd.Uniform(low=0, high=2), name="x")
        y = yield J.Prior(tfpd.Normal(loc=2, scale=x), name="y")
        return x + y

    # z = x + y ~ N(2 + x, x^2) given x: E[exp(-z^2)] = exp(-(2 + x)^2 / (1 + 2 x^2)) / sqrt(1 + 2 x^2)
    f = lambda x: 0.5 * math.exp(-(2 + x) ** 2 / (1 + 2 * x * x)) / math.sqrt(1 + 2 * x * x)
    return J.Model(prior_model, lambda z: -z ** 2), math.log(integrate.quad(f, 0, 2, epsabs=1e-13)[0]), 2000


def _plateau():
    def prior_model():
        x = yield J.Prior(tfpd.Uniform(low=0, high=1))
        return x

    return J.Model(prior_model, lambda x: 0.0), 0.0, 1000


def _basic_mvn():
    import torch
    D = 8
    prior_mu, data_mu = 15 * np.ones(D), np.zeros(D)
    data_cov = np.where(np.eye(D) == 1, 1.0, 0.99)
    log_Z = multivariate_normal(prior_mu, np.eye(D) + data_cov).logpdf(data_mu)
    dist = torch.distributions.MultivariateNormal(
        torch.zeros(D, dtype=torch.float64, device="cuda"),
        scale_tril=torch.linalg.cholesky(torch.tensor(data_cov, dtype=torch.float64, device="cuda")))

    def prior_model():
        x = yield J.Prior(tfpd.MultivariateNormalTriL(loc=prior_mu, scale_tril=np.eye(D)), name="x")
        return x

    return J.Model(prior_model, dist.log_prob), float(log_Z), 100000


@pytest.mark.parametrize("name", ["basic", "basic_with_obj", "basic2", "basic3", "plateau", "basic_mvn"])
def test_reference_fixture(name):
    model, log_Z_true, max_samples = globals()["_" + name]()
    ns = J.NestedSampler(model=model, max_samples=max_samples)
    reason, state = ns(42)
    res = ns.to_results(reason, state)
    z = J.sample_evidence(42, res.num_live_points_per_sample, res.log_L_samples, S=1000)
    z = z[(z > np.percentile(z, 5)) & (z < np.percentile(z, 95))]
    m, sd = float(np.mean(z)), float(np.std(z))
    say(name, "log Z", res.log_Z_mean, "+-", res.log_Z_uncert, "ensemble", m, sd, "true", log_Z_true,
        "samples", res.total_num_samples, "reason", reason)
    assert np.isfinite(res.log_Z_mean) and np.isfinite(res.log_Z_uncert)
    np.testing.assert_allclose(m, log_Z_true, atol=3.0 * res.log_Z_uncert)
    np.testing.assert_allclose(res.log_Z_mean, m, atol=3.0 * res.log_Z_uncert)
    np.testing.assert_allclose(res.log_Z_uncert, sd, atol=math.sqrt(res.log_Z_uncert ** 2 + sd ** 2))
