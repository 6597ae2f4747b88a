"""The one-off FP64 preparation of the multivariate-normal likelihood on the GPU: the Cholesky factor of
a covariance (hipns_cholesky_f64) against LAPACK through numpy, and the sampling kernel's parameter block
(hipns_mvn_prepare_f64: W = L^-1 S, b = L^-1 (loc - mean), c) against FP64 solves on the host and
against scipy's multivariate-normal log density."""
import numpy as np
import pytest
from scipy import stats
from scipy.linalg import solve_triangular
from scipy.special import ndtri

import jaxns_amd as J
from jaxns_amd import distributions as tfpd
from jaxns_amd.likelihoods import device_cholesky, mvn_prepare_device

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _spd(D, seed):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(D, D))
    return A @ A.T / D + 0.05 * np.eye(D)


@pytest.mark.parametrize("D", [1, 3, 10, 32, 70, 128])
def test_cholesky_f64_vs_lapack(D):
    say(f"cholesky D={D}")
    A = _spd(D, D)
    L = device_cholesky(A)
    ref = np.linalg.cholesky(A)
    assert np.all(np.triu(L, 1) == 0.0)
    np.testing.assert_allclose(L, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    np.testing.assert_allclose(L @ L.T, A, rtol=0, atol=1e-13 * np.abs(A).max())


def test_cholesky_f64_reports_indefinite():
    A = _spd(6, 1)
    A[4, 4] = -1.0  # the leading minor of order 5 is not positive definite
    with pytest.raises(np.linalg.LinAlgError, match="leading minor 5"):
        device_cholesky(A)


@pytest.mark.parametrize("prior,D", [("normal", 10), ("tril", 5), ("normal", 70)])
def test_mvn_prepare_matches_fp64_solves(prior, D):
    """The FP32 block holds float32 of W and b formed by FP64 solves against the device factor (1 ulp),
    and c - |W q(u) + b|^2 / 2 equals scipy's log N(x(u); mean, cov) (1e-5 relative, the FP32 block)."""
    say(f"mvn prepare {prior} D={D}")
    rng = np.random.default_rng(D)
    cov = _spd(D, D + 100)
    mean = rng.normal(size=D)
    if prior == "normal":
        loc, scale = rng.normal(size=D), rng.uniform(0.5, 2, size=D)

        def prior_model():
            x = yield J.Prior(tfpd.Normal(loc, scale), name="x")
            return x
    else:
        loc = rng.normal(size=D)
        Lp = np.linalg.cholesky(_spd(D, 7))

        def prior_model():
            x = yield J.Prior(tfpd.MultivariateNormalTriL(loc=loc, scale_tril=Lp), name="x")
            return x

    lik = J.MultivariateNormalLikelihood(mean, cov=cov)
    model = J.Model(prior_model, lik)
    dps = 32 * ((D + 31) // 32)
    prm, c, L = mvn_prepare_device(lik, model, dps)
    prm = prm.cpu().numpy()
    np.testing.assert_allclose(L, np.linalg.cholesky(cov), rtol=0, atol=1e-12)
    qkind, ploc, S = model.prior_program()
    W = solve_triangular(L, S, lower=True)
    b = solve_triangular(L, ploc - mean, lower=True)
    assert np.array_equal(prm[:D], qkind.astype(np.float32))
    ulp = np.spacing(np.abs(np.concatenate([b, W.ravel()])).astype(np.float32))
    got = np.concatenate([prm[dps:dps + D], prm[2 * dps:].reshape(dps, dps)[:D, :D].ravel()])
    assert np.all(np.abs(got.astype(np.float64) - np.concatenate([b, W.ravel()])) <= ulp)
    assert np.all(prm[dps + D:2 * dps] == 0) and np.all(prm[2 * dps:].reshape(dps, dps)[D:] == 0)
    assert c == pytest.approx(-0.5 * D * np.log(2 * np.pi) - np.sum(np.log(np.diag(L))), rel=1e-14)
    U = rng.uniform(0.05, 0.95, size=(40, D))
    z = np.where(qkind == 1, ndtri(U), U)
    r = z @ prm[2 * dps:].reshape(dps, dps)[:D, :D].astype(np.float64).T + prm[dps:dps + D]
    X = ploc + z @ S.T
    want = stats.multivariate_normal(mean, cov).logpdf(X)
    np.testing.assert_allclose(c - 0.5 * np.sum(r * r, axis=1), want, rtol=1e-5)
    # the same likelihood given its factor (scale_tril) runs only the solves
    lik2 = J.MultivariateNormalLikelihood(mean, scale_tril=L)
    prm2, c2 = lik2.device_params(J.Model(prior_model, lik2), dps)
    assert np.array_equal(prm2.cpu().numpy(), prm) and c2 == c
