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
    """log N(x; mean, cov) (tests/conftest.py:219-224).

    Given a covariance, its Cholesky factor is formed once on the GPU in FP64 (hipns_cholesky_f64);
    the sampling kernel's parameters (W = L^-1 S, b = L^-1 (loc - mean) for the model's affine prior)
    come from the same factor by FP64 solves on the GPU (hipns_mvn_prepare_f64)."""

    family = FAM_MVN

    def __init__(self, mean, cov=None, scale_tril=None):
        self.mean = np.atleast_1d(np.asarray(mean, dtype=np.float64))
        self.D = self.mean.size
        if cov is None and scale_tril is None:
            raise ValueError("give cov or scale_tril")
        self.cov = None if cov is None else np.asarray(cov, dtype=np.float64).reshape(self.D, self.D)
        self._scale_tril = None if scale_tril is None else np.tril(np.asarray(scale_tril, dtype=np.float64))
        self._prepared = {}

    @property
    def scale_tril(self) -> np.ndarray:
        if self._scale_tril is None:
            self._scale_tril = device_cholesky(self.cov)
        return self._scale_tril

    @property
    def log_const(self) -> float:
        return float(-0.5 * self.D * np.log(2 * np.pi) - np.sum(np.log(np.diag(self.scale_tril))))

    def log_prob_numpy(self, X):
        from scipy.linalg import solve_triangular
        X = np.atleast_2d(np.asarray(X, dtype=np.float64))
        Y = solve_triangular(self.scale_tril, (X - self.mean).T, lower=True)
        return self.log_const - 0.5 * np.sum(Y * Y, axis=0)

    def torch_log_prob(self, x):
        """The same density in torch (float64, batched [B, D]) for the stepwise path."""
        import torch
        L = torch.as_tensor(self.scale_tril, dtype=torch.float64, device=x.device)
        mu = torch.as_tensor(self.mean, dtype=torch.float64, device=x.device)
        z = torch.linalg.solve_triangular(L, (x - mu).T, upper=False)
        return self.log_const - 0.5 * (z * z).sum(0)

    def device_params(self, model, dps):
        """(FP32 parameter block on the GPU, log_const), formed once per (model, dps) by
        hipns_mvn_prepare_f64."""
        D = model.U_ndims
        if D != self.D:
            raise ValueError(f"likelihood dimension {self.D} != model U dimension {D}")
        key = (id(model), dps)
        if key not in self._prepared:
            prm, c, L = mvn_prepare_device(self, model, dps)
            if self._scale_tril is None:
                self._scale_tril = L
            self._prepared[key] = (model, prm, c)
        _, prm, c = self._prepared[key]
        return prm, c


def device_cholesky(A: np.ndarray) -> np.ndarray:
    """Lower Cholesky factor of A in FP64 on the GPU (hipns_cholesky_f64)."""
    import torch
    from ._lib import check, lib, ptr, stream_ptr
    A = np.ascontiguousarray(A, dtype=np.float64)
    D = A.shape[0]
    Ad = torch.from_numpy(A).cuda()
    Ld = torch.empty_like(Ad)
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    check(lib().hipns_cholesky_f64(ptr(Ad), D, ptr(Ld), ptr(info), stream_ptr()), "hipns_cholesky_f64")
    if int(info.item()):
        raise np.linalg.LinAlgError(f"covariance is not positive definite (leading minor {int(info.item())})")
    return Ld.cpu().numpy()


def mvn_prepare_device(lik: "MultivariateNormalLikelihood", model, dps: int):
    """hipns_mvn_prepare_f64 for the model's affine prior: (params [2 dps + dps^2] float32 on the GPU,
    log_const, L as a host array)."""
    import torch
    from ._lib import check, lib, ptr, stream_ptr
    D = lik.D
    qkind, loc, S = model.prior_program()
    f64 = dict(dtype=torch.float64, device="cuda")
    mu = torch.as_tensor(lik.mean, **f64)
    locd = torch.as_tensor(np.ascontiguousarray(loc), **f64)
    Sd = torch.as_tensor(np.ascontiguousarray(S), **f64)
    qk = torch.as_tensor(np.ascontiguousarray(qkind, dtype=np.int32), device="cuda")
    if lik.cov is not None and lik._scale_tril is None:
        cov, Ld = torch.as_tensor(lik.cov, **f64), torch.empty((D, D), **f64)
    else:
        cov, Ld = None, torch.as_tensor(np.ascontiguousarray(lik.scale_tril), **f64)
    prm = torch.empty(2 * dps + dps * dps, dtype=torch.float32, device="cuda")
    c = torch.empty(1, **f64)
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    ws = torch.empty(int(lib().hipns_mvn_prepare_workspace_size(D)), dtype=torch.uint8, device="cuda")
    check(lib().hipns_mvn_prepare_f64(None if cov is None else ptr(cov), ptr(mu), ptr(locd), ptr(Sd), ptr(qk), D, dps,
                                      ptr(Ld), ptr(prm), ptr(c), ptr(info), ptr(ws), ws.numel(), stream_ptr()),
          "hipns_mvn_prepare_f64")
    if int(info.item()):
        raise np.linalg.LinAlgError(f"covariance is not positive definite (leading minor {int(info.item())})")
    return prm, float(c.item()), Ld.cpu().numpy()


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

    def torch_log_prob(self, x):
        """The same density in torch (float64, batched [B, D]) for the stepwise path."""
        import torch
        m = torch.as_tensor(self.means, dtype=torch.float64, device=x.device)
        sg = torch.as_tensor(self.sigmas, dtype=torch.float64, device=x.device)
        lc = torch.as_tensor(self.logc, dtype=torch.float64, device=x.device)
        d2 = ((x[:, None, :] - m[None]) ** 2).sum(-1)
        return torch.logsumexp(lc[None] - 0.5 * d2 / sg[None] ** 2, dim=1)

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


def _bf16_rne(a: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest, ties to even); returns the 16-bit patterns."""
    b = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)
    r = (b + np.uint32(0x7FFF) + ((b >> np.uint32(16)) & np.uint32(1))) >> np.uint32(16)
    return r.astype(np.uint16)


def _bf16_to_f32(h: np.ndarray) -> np.ndarray:
    return (h.astype(np.uint32) << np.uint32(16)).view(np.float32)


class RegressionData:
    """Design matrix and targets shared by every run of a batch.

    The device evaluates r = r0 - X_s (w_s - w0_s) with r0 = y - X w0 formed once in FP64, where
    X_s = X / c has unit-RMS columns (per-column scales c) and w_s = c (.) w.  Near the posterior
    w - w0 is small, so the BF16x3 product carries only the small correction and the large
    residual part r0 stays exact to FP32 rounding.
    """

    def __init__(self, X, y):
        self.X = np.ascontiguousarray(X, dtype=np.float64)
        self.y = np.ascontiguousarray(y, dtype=np.float64)
        self.N, self.D = self.X.shape
        if self.y.shape != (self.N,):
            raise ValueError("y must have one entry per row of X")
        c = np.sqrt(np.mean(self.X ** 2, axis=0))
        self.col_scale = np.where(c > 0, c, 1.0)
        self.w0 = np.linalg.lstsq(self.X, self.y, rcond=None)[0]
        self._packed = {}

    def packed(self, dps: int):
        """Shared device block as float32 words: [r0: npad][X hi][X lo][y: npad][X], with the
        column-scaled X split into bf16 hi + lo parts and also kept in FP32 (the plain form), each stored
        as `panels` column panels of 16 ([panel][npad][16]); returns (block, npad, panels)."""
        if dps in self._packed:
            return self._packed[dps]
        npad = -(-self.N // 32) * 32
        # kernel instances read 32, 64 or 128 columns (the dimension class of D), zero beyond D
        panels = 2 if self.D <= 32 else (4 if self.D <= 64 else 8)
        r0 = np.zeros(npad, np.float32)
        r0[:self.N] = self.y - self.X @ self.w0
        Xs = np.zeros((npad, 16 * panels), np.float32)
        Xs[:self.N, :self.D] = self.X / self.col_scale
        hi = _bf16_rne(Xs)
        lo = _bf16_rne(Xs - _bf16_to_f32(hi))
        to_panels = lambda a: np.ascontiguousarray(a.reshape(npad, panels, 16).transpose(1, 0, 2)).ravel()
        yp = np.zeros(npad, np.float32)
        yp[:self.N] = self.y
        blk = np.concatenate([r0, to_panels(hi).view(np.float32), to_panels(lo).view(np.float32), yp,
                              to_panels(Xs)])
        self._packed[dps] = (blk, npad, panels)
        return self._packed[dps]


class LinearRegressionLikelihood(LikelihoodFamily):
    """-N/2 log(2 pi sigma^2) - |y - X w|^2 / (2 sigma^2) over the weights w (the prior's output).

    Runs of one batch share the RegressionData object (pass data=...), and may differ in sigma and
    in their priors.  Device layout per run: [q-kind: dps][p: dps][s: dps][w0: dps][1/(2 sig^2) as FP64, 0, 0]
    in column-scaled coordinates, then p, s and w0 as FP64 [3][dps] (the refinement form), plus the
    shared block of RegressionData.packed.
    """

    family = FAM_LINREG

    def __init__(self, X=None, y=None, sigma=1.0, data: "RegressionData" = None):
        self.data = data if data is not None else RegressionData(X, y)
        self.X, self.y = self.data.X, self.data.y
        self.sigma = float(sigma)
        self.N, self.D = self.X.shape
        self.log_const = -0.5 * self.N * np.log(2 * np.pi * self.sigma ** 2)

    def with_sigma(self, sigma) -> "LinearRegressionLikelihood":
        return LinearRegressionLikelihood(sigma=sigma, data=self.data)

    def log_prob_numpy(self, W):
        W = np.atleast_2d(np.asarray(W, dtype=np.float64))
        R = self.y[None, :] - W @ self.X.T
        return self.log_const - 0.5 * np.sum(R * R, axis=1) / self.sigma ** 2

    def torch_log_prob(self, w):
        """The same density in torch (float64, batched [B, D]) for the stepwise path."""
        import torch
        X = torch.as_tensor(self.X, dtype=torch.float64, device=w.device)
        y = torch.as_tensor(self.y, dtype=torch.float64, device=w.device)
        R = y[None, :] - w @ X.T
        return self.log_const - 0.5 * (R * R).sum(1) / self.sigma ** 2

    def device_params(self, model, dps):
        D = model.U_ndims
        if D != self.D:
            raise ValueError(f"likelihood dimension {self.D} != model U dimension {D}")
        if not model.is_diagonal_prior:
            raise ValueError("the regression family needs independent (diagonal) priors on the weights")
        qkind, loc, S = model.prior_program()
        c = self.data.col_scale
        prm = np.zeros(4 * dps + 4, np.float32)
        prm[:D] = qkind
        prm[dps:dps + D] = loc * c
        prm[2 * dps:2 * dps + D] = np.diag(S) * c
        prm[3 * dps:3 * dps + D] = self.data.w0 * c
        prm[4 * dps:4 * dps + 2] = np.array([0.5 / self.sigma ** 2], np.float64).view(np.float32)
        # p, s and w0 again in FP64 for the refinement form (lr_terms 5)
        f64 = np.zeros((3, dps), np.float64)
        f64[0, :D] = loc * c
        f64[1, :D] = np.diag(S) * c
        f64[2, :D] = self.data.w0 * c
        return np.concatenate([prm, f64.ravel().view(np.float32)]), float(self.log_const)

    def log_evidence_normal_prior(self, tau) -> float:
        """Analytic log Z for w ~ N(0, tau^2 I): log N(y; 0, sigma^2 I + tau^2 X X^T) (Woodbury form)."""
        s2, t2 = self.sigma ** 2, float(tau) ** 2
        A = np.eye(self.D) / t2 + self.X.T @ self.X / s2
        Lc = np.linalg.cholesky(A)
        b = self.X.T @ self.y / s2
        m = np.linalg.solve(A, b)
        logdet = 2 * np.sum(np.log(np.diag(Lc))) + self.D * np.log(t2) + self.N * np.log(s2)
        quad = self.y @ self.y / s2 - b @ m
        return float(-0.5 * (self.N * np.log(2 * np.pi) + logdet + quad))

    def grad_log_evidence_normal_prior(self, tau):
        """Analytic d log Z / d(log sigma, log tau) for w ~ N(0, tau^2 I)."""
        s2, t2 = self.sigma ** 2, float(tau) ** 2
        A = np.eye(self.D) / t2 + self.X.T @ self.X / s2
        Ainv = np.linalg.inv(A)
        m = Ainv @ (self.X.T @ self.y / s2)
        r = self.y - self.X @ m
        # posterior moments: E|y - Xw|^2 = |r|^2 + tr(X Ainv X^T); E|w|^2 = |m|^2 + tr(Ainv)
        e_rr = r @ r + np.trace(self.X @ Ainv @ self.X.T)
        e_ww = m @ m + np.trace(Ainv)
        return np.array([-self.N + e_rr / s2, -self.D + e_ww / t2])


class TorchLikelihood:
    """A likelihood given as a batched torch callable, evaluated between sampler steps (the general
    form).  ``fn`` receives what the prior model returns (a tuple is unpacked into arguments), as
    float64 device tensors batched along a leading axis, and returns log-likelihoods of shape [B]
    (a scalar is broadcast); NaN is read as -inf."""

    family = None

    def __init__(self, fn, model):
        import torch
        self.fn = fn
        self.model = model
        self.torch = torch

    def __call__(self, U):
        t = self.torch
        L = self.fn(*self.model.likelihood_inputs_torch(U))
        L = t.as_tensor(L, dtype=t.float64, device=U.device)
        L = L.expand(U.shape[0]) if L.dim() == 0 else L.reshape(-1)
        return t.where(t.isnan(L), t.full_like(L, -float("inf")), L)
