"""Prior / Model surface of jaxns (framework/prior.py:65, framework/model.py:24).

A prior model is a generator function that yields ``Prior(distribution)``
objects and returns the likelihood's inputs, exactly as in jaxns::

    def prior_model():
        x = yield Prior(tfpd.Normal(loc=jnp.zeros(D), scale=jnp.ones(D)), name='x')
        return x

The generator is executed once on the host with placeholder values to record
the prior program: for each prior its U-space width, its per-dimension inverse
CDF q (ndtri for Gaussian families, identity for uniform ones) and the affine
map x = loc + S q(u) that TFP's quantile + bijector apply
(framework/wrapped_tfp_distribution.py:77-84).  The device kernels evaluate
that program inside the sampling loop.
"""
from __future__ import annotations

import inspect
from typing import Any, Callable, List, Optional

import numpy as np
from scipy.special import ndtri

from . import distributions as tfpd


class InvalidPriorName(Exception):
    def __init__(self, name: Optional[str] = None):
        super().__init__(f"Prior name {name} already taken by another prior.")


class Prior:
    """framework/prior.py:65 -- a distribution with an optional name."""

    def __init__(self, dist_or_value, name: Optional[str] = None):
        if isinstance(dist_or_value, tfpd.Distribution):
            self.dist = dist_or_value
            self._type = "dist"
        else:
            self.value = np.asarray(dist_or_value)
            self._type = "value"
        self.name = name

    @property
    def base_ndims(self) -> int:
        return 0 if self._type == "value" else self.dist.num_dims

    def forward(self, U: np.ndarray) -> np.ndarray:
        if self._type == "value":
            return self.value
        return self.dist.forward(U)

    def __repr__(self):
        return f"Prior({self.dist if self._type == 'dist' else self.value}, name={self.name})"


def _run_generator(prior_model: Callable, U: np.ndarray):
    """Drive the prior generator with U (framework/ops.py:240-267); returns (priors, outputs, X list)."""
    gen = prior_model()
    if not inspect.isgenerator(gen):
        raise ValueError("prior_model must be a generator function yielding Prior objects")
    priors: List[Prior] = []
    xs = []
    resp = None
    idx = 0
    names = set()
    while True:
        try:
            p = gen.send(resp)
        except StopIteration as e:
            out = e.value
            break
        if not isinstance(p, Prior):
            raise ValueError(f"prior_model yielded {type(p)}; expected Prior")
        if p.name is not None:
            if p.name in names:
                raise InvalidPriorName(p.name)
            names.add(p.name)
        w = p.base_ndims
        x = p.forward(U[idx:idx + w])
        idx += w
        priors.append(p)
        xs.append(x)
        resp = x
    return priors, out, xs


class Model:
    """framework/model.py:24 -- Model(prior_model, log_likelihood).

    ``log_likelihood`` is either one of the fused likelihood families of
    :mod:`jaxns_amd.likelihoods` (evaluated inside the sampling kernel) or any
    callable taking the prior model's outputs (evaluated with numpy for host
    checks only).
    """

    def __init__(self, prior_model: Callable, log_likelihood: Any, params: Optional[Any] = None):
        self.prior_model = prior_model
        self.log_likelihood = log_likelihood
        self._params = params
        priors, out, xs = _run_generator(prior_model, np.full(4096, 0.5))
        self.priors = [p for p in priors if p._type == "dist"]
        self.U_ndims = int(sum(p.base_ndims for p in priors))
        if self.U_ndims < 1:
            raise ValueError("model has no random variables")
        self._check_output_is_concatenation(out, xs)

    def __repr__(self):
        return f"Model(U_ndims={self.U_ndims}, num_params=0)"

    def _check_output_is_concatenation(self, out, xs):
        flat = np.concatenate([np.ravel(x) for x in xs]) if xs else np.zeros(0)
        outs = out if isinstance(out, tuple) else (out,)
        try:
            got = np.concatenate([np.ravel(np.asarray(o, dtype=float)) for o in outs])
        except Exception:
            got = None
        self._output_is_concat = got is not None and got.shape == flat.shape and np.allclose(got, flat)
        # a prior whose parameters depend on earlier prior values (e.g. Normal(loc, scale=x)) has no
        # fixed affine program: detect it by running the generator at two other points
        self.dynamic_prior = False
        forms = []
        for v in (0.3, 0.7):
            pr, _, _ = _run_generator(self.prior_model, np.full(self.U_ndims, v))
            forms.append([p.dist.affine_form() for p in pr if p._type == "dist"])
        for fa, fb in zip(*forms):
            if any(not np.allclose(np.asarray(x), np.asarray(y)) for x, y in zip(fa, fb)):
                self.dynamic_prior = True

    def likelihood_inputs_torch(self, U):
        """The prior model's return value for a torch batch U [B, >= U_ndims] in float64, computed by
        driving the generator with batched prior values (framework/ops.py:240-267, prepare_input):
        each yielded prior maps its U slice through its quantile and bijector with the parameters it
        was given, which may themselves be earlier prior values.  Returned as a tuple of likelihood
        arguments (a tuple return value is unpacked, as in jaxns)."""
        import torch
        Bn = U.shape[0]
        u64 = U[:, :self.U_ndims].double()
        gen = self.prior_model()
        resp = None
        i = o = 0
        while True:
            try:
                p = gen.send(resp)
            except StopIteration as e:
                out = e.value
                break
            if p._type == "value":
                resp = torch.as_tensor(p.value, dtype=torch.float64, device=U.device)
                continue
            ref = self.priors[i]
            w = ref.base_ndims
            shp = tuple(getattr(ref.dist, "event_shape", (w,)))
            resp = p.dist.forward_torch(u64[:, o:o + w].reshape((Bn,) + shp))
            i += 1
            o += w
        return out if isinstance(out, tuple) else (out,)

    # ------------------------------------------------------------------ program
    def prior_program(self):
        """(qkind [D] int, loc [D], dense S [D, D]) with X = loc + S q(U)."""
        D = self.U_ndims
        qkind = np.zeros(D, np.int32)
        loc = np.zeros(D)
        S = np.zeros((D, D))
        o = 0
        for p in self.priors:
            k, l, s = p.dist.affine_form()
            w = p.base_ndims
            qkind[o:o + w] = k
            loc[o:o + w] = l
            S[o:o + w, o:o + w] = s
            o += w
        return qkind, loc, S

    @property
    def is_diagonal_prior(self) -> bool:
        _, _, S = self.prior_program()
        return np.allclose(S, np.diag(np.diag(S)))

    def transform(self, U: np.ndarray):
        """Model.transform (framework/model.py:155-159): U -> X (numpy, float64), dict by prior name."""
        U = np.atleast_2d(np.asarray(U, dtype=np.float64))
        out = {}
        o = 0
        for i, p in enumerate(self.priors):
            w = p.base_ndims
            out[p.name if p.name is not None else f"_{i}"] = p.dist.forward(U[:, o:o + w])
            o += w
        return out

    def split_transformed(self, X: np.ndarray):
        """The dict of Model.transform from X = loc + S q(U) formed elsewhere (rows = samples): each
        prior's columns under its name, a scalar prior's as a vector."""
        out = {}
        o = 0
        for i, p in enumerate(self.priors):
            w = p.base_ndims
            x = X[:, o:o + w]
            if getattr(p.dist, "event_shape", None) == ():
                x = x[:, 0]
            out[p.name if p.name is not None else f"_{i}"] = x
            o += w
        return out

    def forward_numpy(self, U: np.ndarray) -> np.ndarray:
        """Model.forward on the host in float64 (framework/model.py:167-176); NaN -> -inf."""
        U = np.atleast_2d(np.asarray(U, dtype=np.float64))
        qkind, loc, S = self.prior_program()
        q = np.where(qkind[None, :] == 1, ndtri(U), U)
        X = loc[None, :] + q @ S.T
        if hasattr(self.log_likelihood, "log_prob_numpy"):
            L = self.log_likelihood.log_prob_numpy(X)
        else:
            L = np.array([float(self.log_likelihood(x)) for x in X])
        return np.where(np.isnan(L), -np.inf, L)

    def log_prob_prior_numpy(self, U: np.ndarray) -> np.ndarray:
        U = np.atleast_2d(np.asarray(U, dtype=np.float64))
        lp = np.zeros(U.shape[0])
        o = 0
        for p in self.priors:
            w = p.base_ndims
            lp += p.dist.log_prob_from_u(U[:, o:o + w])
            o += w
        return lp

    def sample_U(self, key, num: int = 1) -> np.ndarray:
        rng = np.random.default_rng(int(np.asarray(key).ravel()[0]))
        return rng.uniform(size=(num, self.U_ndims))


PriorModelGen = Any
