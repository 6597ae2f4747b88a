"""Likelihood families fused into the sampling kernels.

A family object is passed as ``Model(prior_model, log_likelihood=<family>)``.
It evaluates on the concatenated prior outputs x (in yield order) and packs
its parameters for the device in the layout the kernels read
(jaxns_amd/csrc/ns_sampler_impl.h):

  MVN : [q-kind: dps][b: dps][W: dps x dps]          r = W q(u) + b, logL = c - |r|^2/2
  GMM : [q-kind: dps][p: dps][s: dps][means: K x dps][1/(2 sig^2): K][log w - D/2 log(2 pi sig^2): K]

The user-function equivalents are the benchmark likelihoods of jaxns'
tests (tests/conftest.py:219-224 log_normal) and the mixture / regression
likelihoods of the benchmark configurations.
"""
from __future__ import annotations

import numpy as np
from scipy.special import logsumexp

from ._lib import FAM_GMM, FAM_LINREG, FAM_MVN, MAX_GMM_K


class LikelihoodFamily:
    family: int

    def log_prob_numpy(self, X: np.ndarray) -> np.ndarray:
        raise NotImplementedError

    def __call__(self, x):
        return self.log_prob_numpy(np.atleast_2d(np.asarray(x, dtype=np.float64)))[0]

    def device_params(self, model, dps: int):
        raise NotImplementedError


class MultivariateNormalLikelihood(LikelihoodFamily):
    """log N(x; mean, cov) (tests/conftest.py:219-224)."""

    family = FAM_MVN

    def __init__(self, mean, cov=None, scale_tril=None):
        self.mean = np.atleast_1d(np.asarray(mean, dtype=np.float64))
        if scale_tril is None:
            cov = np.asarray(cov, dtype=np.float64)
            scale_tril = np.linalg.cholesky(cov)
        self.scale_tril = np.tril(np.asarray(scale_tril, dtype=np.float64))
        self.D = self.mean.size
        self.log_const = -0.5 * self.D * np.log(2 * np.pi) - np.sum(np.log(np.diag(self.scale_tril)))

    def log_prob_numpy(self, X):
        from scipy.linalg import solve_triangular
        X = np.atleast_2d(np.asarray(X, dtype=np.float64))
        Y = solve_triangular(self.scale_tril, (X - self.mean).T, lower=True)
        return self.log_const - 0.5 * np.sum(Y * Y, axis=0)

    def device_params(self, model, dps):
        from scipy.linalg import solve_triangular
        D = model.U_ndims
        if D != self.D:
            raise ValueError(f"likelihood dimension {self.D} != model U dimension {D}")
        qkind, loc, S = model.prior_program()
        W = solve_triangular(self.scale_tril, S, lower=True)
        b = solve_triangular(self.scale_tril, loc - self.mean, lower=True)
        prm = np.zeros(2 * dps + dps * dps, np.float32)
        prm[:D] = qkind
        prm[dps:dps + D] = b
        Wp = np.zeros((dps, dps))
        Wp[:D, :D] = W
        prm[2 * dps:] = Wp.ravel()
        return prm, float(self.log_const)


class GaussianMixtureLikelihood(LikelihoodFamily):
    """log sum_k w_k N(x; m_k, sigma_k^2 I) with isotropic components."""

    family = FAM_GMM

    def __init__(self, means, sigmas, weights=None):
        self.means = np.atleast_2d(np.asarray(means, dtype=np.float64))
        self.sigmas = np.atleast_1d(np.asarray(sigmas, dtype=np.float64))
        self.K, self.D = self.means.shape
        if self.K > MAX_GMM_K:
            raise ValueError(f"at most {MAX_GMM_K} components")
        w = np.full(self.K, 1.0 / self.K) if weights is None else np.asarray(weights, dtype=np.float64)
        self.log_w = np.log(w / w.sum())
        self.logc = self.log_w - 0.5 * self.D * np.log(2 * np.pi * self.sigmas ** 2)

    def log_prob_numpy(self, X):
        X = np.atleast_2d(np.asarray(X, dtype=np.float64))
        d2 = np.sum((X[:, None, :] - self.means[None]) ** 2, axis=-1)
        return logsumexp(self.logc[None] - 0.5 * d2 / self.sigmas[None] ** 2, axis=1)

    def device_params(self, model, dps):
        D = model.U_ndims
        if D != self.D:
            raise ValueError(f"likelihood dimension {self.D} != model U dimension {D}")
        if not model.is_diagonal_prior:
            raise ValueError("the mixture family needs independent (diagonal) priors")
        qkind, loc, S = model.prior_program()
        K = self.K
        prm = np.zeros((3 + K) * dps + 2 * K, np.float32)
        prm[:D] = qkind
        prm[dps:dps + D] = loc
        prm[2 * dps:2 * dps + D] = np.diag(S)
        for k in range(K):
            prm[(3 + k) * dps:(3 + k) * dps + D] = self.means[k]
        prm[(3 + K) * dps:(3 + K) * dps + K] = 0.5 / self.sigmas ** 2
        prm[(3 + K) * dps + K:] = self.logc
        return prm, 0.0

    def log_evidence_uniform_prior(self, low, high):
        """Analytic log Z under a uniform box prior [low, high]^D."""
        from scipy.special import log_ndtr
        low = np.broadcast_to(np.asarray(low, float), (self.D,))
        high = np.broadcast_to(np.asarray(high, float), (self.D,))
        terms = []
        for k in range(self.K):
            a = (low - self.means[k]) / self.sigmas[k]
            b = (high - self.means[k]) / self.sigmas[k]
            # log(Phi(b) - Phi(a)) stably
            lb, la = log_ndtr(b), log_ndtr(a)
            terms.append(self.log_w[k] + np.sum(lb + np.log1p(-np.exp(la - lb))))
        return float(logsumexp(terms) - np.sum(np.log(high - low)))


class LinearRegressionLikelihood(LikelihoodFamily):
    """-N/2 log(2 pi sigma^2) - |y - X w|^2 / (2 sigma^2)."""

    family = FAM_LINREG

    def __init__(self, X, y, sigma):
        self.X = np.asarray(X, dtype=np.float64)
        self.y = np.asarray(y, dtype=np.float64)
        self.sigma = float(sigma)
        self.N, self.D = self.X.shape

    def log_prob_numpy(self, W):
        W = np.atleast_2d(np.asarray(W, dtype=np.float64))
        R = self.y[None, :] - W @ self.X.T
        return -0.5 * self.N * np.log(2 * np.pi * self.sigma ** 2) - 0.5 * np.sum(R * R, axis=1) / self.sigma ** 2
