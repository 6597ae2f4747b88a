"""Evidence maximisation for the linear-Gaussian family with a hand-written adjoint.

Mirrors jaxns' EvidenceMaximisation (experimental/evidence_maximisation.py:52-300):
the E-step is a nested-sampling run at the current hyperparameters, the M-step
maximises

    log Z(theta) = logsumexp_i [ logL(U_i; theta) + log_weight_i ],
    log_weight_i = log_dp_i - logL_i + log Z          (evidence_maximisation.py:213)

over the E-step's samples U_i (fixed in the unit cube).  jaxns differentiates
this with jax; here the gradient comes from a hand-written adjoint of the batched
likelihood behind a torch.autograd.Function.  For Bayesian linear regression with
theta = (log sigma, log tau), w = tau q(U) (prior N(0, tau^2 I)) and r = y - X w:

    d logL / d log sigma = -N + |r|^2 / sigma^2
    d logL / d log tau   = w^T X^T r / sigma^2 = (y^T r - |r|^2) / sigma^2

The device kernel (hipns_linreg_stats) returns logL, |r|^2 and y^T r per sample
from the same BF16x3 MFMA residual GEMM the sampler uses, so the objective and its
gradient are consistent.  autograd then yields d log Z / d theta as the
posterior-weighted mean of d logL / d theta.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math
from typing import Any, Dict, Optional, Tuple

import numpy as np

from . import _lib
from .framework import Model, Prior
from .likelihoods import LinearRegressionLikelihood, RegressionData
from .types import NestedSamplerResults, TerminationCondition

__all__ = ["LinearRegressionEvidenceModel", "LinRegAdjointStats", "LinRegLogLikelihood", "log_evidence",
           "EvidenceMaximisation"]


class LinearRegressionEvidenceModel:
    """Family of models y ~ N(X w, sigma^2 I), w ~ N(0, tau^2 I) indexed by theta = (log sigma, log tau)."""

    def __init__(self, data: RegressionData, params=(0.0, 0.0)):
        self.data = data
        self.params = np.asarray(params, dtype=np.float64)

    def __call__(self, params=None) -> Model:
        from . import distributions as tfpd
        ls, lt = np.asarray(self.params if params is None else params, dtype=np.float64)
        D = self.data.D
        tau = float(np.exp(lt))

        def prior_model():
            w = yield Prior(tfpd.Normal(loc=np.zeros(D), scale=np.full(D, tau)), name="w")
            return w

        return Model(prior_model, LinearRegressionLikelihood(sigma=float(np.exp(ls)), data=self.data))

    def log_Z_analytic(self, params) -> float:
        ls, lt = params
        return LinearRegressionLikelihood(sigma=math.exp(ls), data=self.data).log_evidence_normal_prior(math.exp(lt))

    def grad_log_Z_analytic(self, params) -> np.ndarray:
        ls, lt = params
        lik = LinearRegressionLikelihood(sigma=math.exp(ls), data=self.data)
        return lik.grad_log_evidence_normal_prior(math.exp(lt))


class LinRegAdjointStats:
    """Device context for a fixed set of unit-cube samples: evaluates (logL, |r|^2, y^T r) per sample
    at any theta.  The shared regression rows are uploaded once; a theta change rewrites only the
    per-run block (4 dps + 4 floats) and the run descriptor."""

    def __init__(self, family: LinearRegressionEvidenceModel, U: np.ndarray, lanes_per_chain: int = 2,
                 lr_terms: int = 3, device: str = "cuda"):
        import torch
        self.family = family
        self.torch = torch
        U = np.asarray(U, dtype=np.float32)
        self.M, self.D = U.shape
        if self.D != family.data.D:
            raise ValueError("U dimension does not match the regression data")
        self.dps = int(math.ceil(self.D / 32) * 32)
        dev = torch.device(device)
        Up = np.full((self.M, self.dps), 0.5, np.float32)
        Up[:, :self.D] = U
        self.U = torch.from_numpy(Up).to(dev)
        blk, npad, panels = family.data.packed(self.dps)
        self.n_run = 4 * self.dps + 4 + 6 * self.dps  # the FP32 block, then p, s, w0 in FP64
        self.data_off = int(-(-self.n_run // 64) * 64)
        prm = np.zeros(self.data_off + blk.size, np.float32)
        prm[self.data_off:] = blk
        self.params = torch.from_numpy(prm).to(dev)
        self.bufs = _lib.Buffers()
        self.bufs.params = _lib.ptr(self.params)
        self.bufs.dps = self.dps
        self.launch = _lib.Launch(_lib.FAM_LINREG, lanes_per_chain, 1, 1, npad, 512, panels, lr_terms, self.data_off)
        self.runs_host = (_lib.RunDesc * 1)()
        self.runs_dev = torch.zeros(C.sizeof(_lib.RunDesc), dtype=torch.uint8, device=dev)
        f64 = torch.float64
        self.logL = torch.empty(self.M, dtype=f64, device=dev)
        self.rss = torch.empty(self.M, dtype=f64, device=dev)
        self.yr = torch.empty(self.M, dtype=f64, device=dev)

    def __call__(self, theta) -> Tuple[Any, Any, Any]:
        torch = self.torch
        th = np.asarray(theta, dtype=np.float64)
        model = self.family(th)
        p, c = model.log_likelihood.device_params(model, self.dps)
        self.params[:self.n_run].copy_(torch.from_numpy(p))
        d = self.runs_host[0]
        d.D = self.D
        d.N = 1
        d.param_off = 0
        d.log_const = c
        self.runs_dev.copy_(torch.frombuffer(bytearray(bytes(self.runs_host)), dtype=torch.uint8))
        _lib.check(_lib.lib().hipns_linreg_stats(_lib.ptr(self.U), self.M, _lib.ptr(self.logL), _lib.ptr(self.rss),
                                                 _lib.ptr(self.yr), _lib.ptr(self.runs_dev), self.runs_host,
                                                 C.byref(self.bufs), C.byref(self.launch), _lib.stream_ptr()))
        return self.logL.clone(), self.rss.clone(), self.yr.clone()


def _autograd_function():
    import torch

    class _LinRegLogLikelihood(torch.autograd.Function):
        """logL(U_i; theta) for the fixed samples of `stats`; backward is the hand-written adjoint."""

        @staticmethod
        def forward(ctx, theta, stats: LinRegAdjointStats):
            th = theta.detach().cpu().numpy()
            logL, rss, yr = stats(th)
            ctx.save_for_backward(rss, yr)
            ctx.sigma2 = math.exp(2.0 * float(th[0]))
            ctx.n_rows = stats.family.data.N
            ctx.theta_device = theta.device
            return logL

        @staticmethod
        def backward(ctx, g):
            rss, yr = ctx.saved_tensors
            s2 = ctx.sigma2
            d_log_sigma = torch.sum(g * (rss / s2 - ctx.n_rows))
            d_log_tau = torch.sum(g * (yr - rss) / s2)
            return torch.stack([d_log_sigma, d_log_tau]).to(ctx.theta_device), None

    return _LinRegLogLikelihood


_FN = None


def LinRegLogLikelihood(theta, stats: LinRegAdjointStats):
    """logL of every sample of `stats` at theta (a float64 torch tensor of shape [2]), differentiable."""
    global _FN
    if _FN is None:
        _FN = _autograd_function()
    return _FN.apply(theta, stats)


def log_evidence(theta, stats: LinRegAdjointStats, log_weights):
    """log Z(theta) = logsumexp(logL(U; theta) + log_weights) (evidence_maximisation.py:151-160)."""
    import torch
    return torch.logsumexp(LinRegLogLikelihood(theta, stats) + log_weights, dim=0)


@dataclasses.dataclass(eq=False)
class EvidenceMaximisation:
    """E-step: nested sampling at theta; M-step: L-BFGS on log Z(theta) over the E-step samples.

    Same arguments and stopping rules as jaxns' EvidenceMaximisation (evidence_maximisation.py:52-80,
    :234-300); `model` is a LinearRegressionEvidenceModel and params are (log sigma, log tau)."""

    model: LinearRegressionEvidenceModel
    ns_kwargs: Optional[Dict[str, Any]] = None
    max_num_epochs: int = 50
    gtol: float = 1e-2
    log_Z_ftol: float = 1.0
    log_Z_atol: float = 1e-4
    batch_size: Optional[int] = None
    termination_cond: Optional[TerminationCondition] = None
    verbose: bool = False
    lr_terms: int = 3

    def __post_init__(self):
        if self.ns_kwargs is None:
            self.ns_kwargs = {}

    def e_step(self, key, params, desc: str = "") -> NestedSamplerResults:
        from .nested_sampler import NestedSampler
        if self.verbose:
            print(f"Running E-step... {desc}", flush=True)
        ns = NestedSampler(model=self.model(params), **self.ns_kwargs)
        reason, state = ns(key, self.termination_cond)
        return ns.to_results(reason, state)

    def m_step_data(self, ns_results: NestedSamplerResults):
        import torch
        lw = (np.asarray(ns_results.log_dp_mean) - np.asarray(ns_results.log_L_samples) +
              float(ns_results.log_Z_mean))
        stats = LinRegAdjointStats(self.model, np.asarray(ns_results.U_samples), lr_terms=self.lr_terms)
        return stats, torch.from_numpy(lw).to(stats.U.device)

    def m_step(self, key, params, ns_results: NestedSamplerResults, desc: str = "") -> Tuple[np.ndarray, float]:
        import torch
        stats, lw = self.m_step_data(ns_results)
        theta = torch.tensor(np.asarray(params, dtype=np.float64), dtype=torch.float64, requires_grad=True,
                             device=stats.U.device)
        opt = torch.optim.LBFGS([theta], lr=1.0, max_iter=20, line_search_fn="strong_wolfe")

        def closure():
            opt.zero_grad()
            loss = -log_evidence(theta, stats, lw)
            loss.backward()
            return loss

        last = theta.detach().cpu().numpy().copy()
        log_Z = float("nan")
        for epoch in range(self.max_num_epochs):
            loss = opt.step(closure)
            log_Z = -float(loss.detach())
            cur = theta.detach().cpu().numpy().copy()
            l_oo = float(np.max(np.abs(cur - last)))
            last = cur
            if self.verbose:
                print(f"{desc}: Epoch {epoch}: log_Z={log_Z}, l_oo={l_oo}", flush=True)
            if l_oo < self.gtol:
                break
        with torch.no_grad():
            log_Z = float(log_evidence(theta, stats, lw))
        return theta.detach().cpu().numpy(), log_Z

    def train(self, num_steps: int = 10, params=None) -> Tuple[NestedSamplerResults, np.ndarray]:
        params = np.asarray(self.model.params if params is None else params, dtype=np.float64)
        log_Z = -np.inf
        ns_results = None
        for step in range(num_steps):
            desc = "Initial run" if ns_results is None else \
                f"Step {step}: log Z = {ns_results.log_Z_mean:.4f} +- {ns_results.log_Z_uncert:.4f}"
            ns_results = self.e_step(key=1000 + step, params=params, desc=desc)
            change = abs(float(ns_results.log_Z_mean) - log_Z)
            if change < self.log_Z_atol or change < self.log_Z_ftol * float(ns_results.log_Z_uncert):
                break
            log_Z = float(ns_results.log_Z_mean)
            params, _ = self.m_step(key=step, params=params, ns_results=ns_results, desc=desc)
        if ns_results is None:
            raise RuntimeError("No results were computed.")
        return ns_results, params
