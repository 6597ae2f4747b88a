"""The distribution set the prior transform supports, with TFP's names.

Each distribution exposes its quantile map in the affine form the device
kernels evaluate: X = loc + S q(U), q = ndtri (Gaussian) or identity
(uniform).  ``forward`` is TFP's ``quantile`` followed by bijectors
(framework/wrapped_tfp_distribution.py:77-84).
"""
from __future__ import annotations

import numpy as np
from scipy.special import ndtri

Q_IDENTITY = 0
Q_NDTRI = 1


def _is_torch(*vals) -> bool:
    return any(type(v).__module__.startswith("torch") for v in vals)


class Distribution:
    num_dims: int

    def affine_form(self):
        raise NotImplementedError

    def forward_torch(self, u):
        """X of a torch batch u of shape (B,) + event_shape (the parameters may be torch tensors
        batched along the same leading axis, as when a prior depends on earlier prior values)."""
        raise NotImplementedError

    def forward(self, U):
        k, loc, S = self.affine_form()
        U = np.asarray(U, dtype=np.float64)
        q = np.where(np.asarray(k) == Q_NDTRI, ndtri(U), U)
        x = loc + q @ S.T
        if U.ndim == 1 and getattr(self, "event_shape", None) == ():
            return x[0]  # a scalar distribution's value is a scalar
        return x

    def log_prob_from_u(self, U):
        raise NotImplementedError


def _vec(a, n=None):
    a = np.atleast_1d(np.asarray(a, dtype=np.float64))
    if n is not None and a.size == 1 and n > 1:
        a = np.full(n, float(a[0]))
    return a


class Normal(Distribution):
    """tfpd.Normal(loc, scale): x = loc + scale * ndtri(u)."""

    def __init__(self, loc=0.0, scale=1.0):
        self._raw = (loc, scale)
        if _is_torch(loc, scale):
            self.num_dims = None
            return
        scalar = np.ndim(loc) == 0 and np.ndim(scale) == 0
        loc = np.atleast_1d(np.asarray(loc, dtype=np.float64))
        scale = np.atleast_1d(np.asarray(scale, dtype=np.float64))
        n = max(loc.size, scale.size)
        self.loc = _vec(loc, n)
        self.scale = _vec(scale, n)
        self.num_dims = n
        self.event_shape = () if scalar else (n,)

    def affine_form(self):
        return np.full(self.num_dims, Q_NDTRI, np.int32), self.loc, np.diag(self.scale)

    def forward_torch(self, u):
        import torch
        loc, scale = (torch.as_tensor(v, dtype=torch.float64, device=u.device) for v in self._raw)
        return loc + scale * torch.special.ndtri(u)

    def log_prob_from_u(self, U):
        z = ndtri(np.asarray(U, dtype=np.float64))
        return np.sum(-0.5 * z * z - 0.5 * np.log(2 * np.pi) - np.log(self.scale), axis=-1)

    def __repr__(self):
        return f"Normal(loc={self.loc}, scale={self.scale})"


class Uniform(Distribution):
    """tfpd.Uniform(low, high): x = low + (high - low) u."""

    def __init__(self, low=0.0, high=1.0):
        self._raw = (low, high)
        if _is_torch(low, high):
            self.num_dims = None
            return
        scalar = np.ndim(low) == 0 and np.ndim(high) == 0
        low = np.atleast_1d(np.asarray(low, dtype=np.float64))
        high = np.atleast_1d(np.asarray(high, dtype=np.float64))
        n = max(low.size, high.size)
        self.low = _vec(low, n)
        self.high = _vec(high, n)
        self.num_dims = n
        self.event_shape = () if scalar else (n,)

    def affine_form(self):
        return np.full(self.num_dims, Q_IDENTITY, np.int32), self.low, np.diag(self.high - self.low)

    def forward_torch(self, u):
        import torch
        low, high = (torch.as_tensor(v, dtype=torch.float64, device=u.device) for v in self._raw)
        return low + (high - low) * u

    def log_prob_from_u(self, U):
        U = np.asarray(U, dtype=np.float64)
        return np.full(U.shape[:-1], -np.sum(np.log(self.high - self.low)))

    def __repr__(self):
        return f"Uniform(low={self.low}, high={self.high})"


class MultivariateNormalTriL(Distribution):
    """tfpd.MultivariateNormalTriL(loc, scale_tril): x = loc + L ndtri(u)."""

    def __init__(self, loc, scale_tril):
        self._raw = (loc, scale_tril)
        if _is_torch(loc, scale_tril):
            self.num_dims = None
            return
        self.loc = np.atleast_1d(np.asarray(loc, dtype=np.float64))
        self.scale_tril = np.tril(np.asarray(scale_tril, dtype=np.float64))
        self.num_dims = self.loc.size
        self.event_shape = (self.num_dims,)
        if self.scale_tril.shape != (self.num_dims, self.num_dims):
            raise ValueError("scale_tril must be [D, D]")

    def affine_form(self):
        return np.full(self.num_dims, Q_NDTRI, np.int32), self.loc, self.scale_tril

    def forward_torch(self, u):
        import torch
        loc, L = (torch.as_tensor(v, dtype=torch.float64, device=u.device) for v in self._raw)
        z = torch.special.ndtri(u)
        return loc + torch.matmul(torch.tril(L), z.unsqueeze(-1)).squeeze(-1)

    def log_prob_from_u(self, U):
        z = ndtri(np.asarray(U, dtype=np.float64))
        return np.sum(-0.5 * z * z - 0.5 * np.log(2 * np.pi), axis=-1) - np.sum(np.log(np.diag(self.scale_tril)))

    def __repr__(self):
        return f"MultivariateNormalTriL(loc={self.loc})"
