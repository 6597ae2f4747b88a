"""Synthetic problems of the benchmark configurations (seeded numpy, FP64 generation)."""
import numpy as np
from scipy.stats import multivariate_normal

import jaxns_amd as J
from jaxns_amd import distributions as tfpd


def mvn_problem(D, seed=0):
    """mu ~ N(0, 0.5^2 I), Sigma = A A^T / D + 0.05 I, prior N(0, I); log Z = log N(mu; 0, I + Sigma)."""
    rng = np.random.default_rng(seed)
    mu = rng.normal(0.0, 0.5, D)
    A = rng.normal(size=(D, D))
    Sigma = A @ A.T / D + 0.05 * np.eye(D)
    log_Z = multivariate_normal(np.zeros(D), np.eye(D) + Sigma).logpdf(mu)

    def prior_model():
        x = yield J.Prior(tfpd.Normal(loc=np.zeros(D), scale=np.ones(D)), name="x")
        return x

    lik = J.MultivariateNormalLikelihood(mu, cov=Sigma)
    return J.Model(prior_model, lik), float(log_Z), mu, Sigma


def gmm_problem(D, seed=0, K=4):
    """4 equal-weight isotropic components, means ~ U[-5,5]^D, sigma ~ U[0.2,1], prior U[-10,10]^D."""
    rng = np.random.default_rng(seed)
    means = rng.uniform(-5, 5, size=(K, D))
    sig = rng.uniform(0.2, 1.0, size=K)

    def prior_model():
        x = yield J.Prior(tfpd.Uniform(low=-10.0 * np.ones(D), high=10.0 * np.ones(D)), name="x")
        return x

    lik = J.GaussianMixtureLikelihood(means, sig)
    return J.Model(prior_model, lik), lik.log_evidence_uniform_prior(-10.0, 10.0), means, sig


def mvn_oracle_spec(D, mu, Sigma):
    import oracle as O
    return O.ModelSpec(O.FAM_MVN, D, O.PRIOR_NORMAL, np.zeros(D), np.ones(D), mvn_mu=mu,
                       mvn_chol=np.linalg.cholesky(Sigma))


def gmm_oracle_spec(D, means, sig):
    import oracle as O
    K = means.shape[0]
    return O.ModelSpec(O.FAM_GMM, D, O.PRIOR_UNIFORM, -10.0 * np.ones(D), 20.0 * np.ones(D), gmm_means=means,
                       gmm_sigma=sig, gmm_logw=np.log(np.full(K, 1.0 / K)))


def linreg_problem(Ndata, D, seed=0, sigma=0.3, tau=1.0, scaled_cols=0.0, data=None):
    """Bayesian linear regression: X ~ N(0,1), w_true ~ N(0,1), y = X w_true + sigma eps, prior w ~ N(0, tau^2 I).
    scaled_cols > 0 multiplies that fraction of the columns of X by 100.  log Z = log N(y; 0, sigma^2 I + tau^2 X X^T)."""
    if data is None:
        rng = np.random.default_rng(seed)
        X = rng.normal(size=(Ndata, D))
        if scaled_cols > 0:
            X[:, rng.permutation(D)[:max(1, int(round(scaled_cols * D)))]] *= 100.0
        w = rng.normal(size=D)
        y = X @ w + sigma * rng.normal(size=Ndata)
        data = J.RegressionData(X, y)
    D = data.D

    def prior_model():
        w = yield J.Prior(tfpd.Normal(loc=np.zeros(D), scale=tau * np.ones(D)), name="w")
        return w

    lik = J.LinearRegressionLikelihood(sigma=sigma, data=data)
    return J.Model(prior_model, lik), lik.log_evidence_normal_prior(tau), lik


def linreg_oracle_spec(lik, tau):
    import oracle as O
    D = lik.D
    return O.ModelSpec(O.FAM_LINREG, D, O.PRIOR_NORMAL, np.zeros(D), tau * np.ones(D), lr_X=lik.X, lr_y=lik.y,
                       lr_sigma=lik.sigma)


def linreg_mvn_oracle_spec(lik, tau):
    """The regression likelihood restated as a Gaussian in w, for long oracle ensembles:
    |y - X w|^2 = rss_min + (w - w_hat)^T X^T X (w - w_hat), so logL(w) is an MVN density with mean
    w_hat and covariance sigma^2 (X^T X)^-1, plus a constant.  Returns (spec, const) with
    logL_linreg(U) = logL_mvn(U) + const; an evaluation costs O(D^2) instead of O(N D)."""
    import oracle as O
    X, y, s2 = lik.X, lik.y, lik.sigma ** 2
    D, N = lik.D, lik.N
    w_hat = np.linalg.solve(X.T @ X, X.T @ y)
    r = y - X @ w_hat
    L = np.linalg.cholesky(s2 * np.linalg.inv(X.T @ X))
    const = (-0.5 * N * np.log(2 * np.pi * s2) - 0.5 * (r @ r) / s2
             + 0.5 * D * np.log(2 * np.pi) + np.sum(np.log(np.diag(L))))
    spec = O.ModelSpec(O.FAM_MVN, D, O.PRIOR_NORMAL, np.zeros(D), tau * np.ones(D), mvn_mu=w_hat, mvn_chol=L)
    return spec, float(const)
