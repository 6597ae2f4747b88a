"""Post-run utilities with jaxns' names (utils.py:284-641, internals/namedtuple_utils.py:34-103).

Results files use jaxns' JSON schema, so that a file written by jaxns loads here and a file written
here loads in jaxns:
  NamedTuple -> {"type": "__namedtuple__", "__class__": "<module>.<name>", "__data__": {field: ...}}
  ndarray    -> {"type": "__ndarray__", "__dtype__": "float64", "__data__": base64 of the raw bytes,
                 "__shape__": [...]}   (jaxns writes jax arrays with "type": "__jax_ndarray__")
  dict / list -> the same container of encoded values; other values as plain JSON.
The result types are written under jaxns' class paths (jaxns.nested_samplers.common.types.*).  On load
a class path is never imported: NestedSamplerResults, TerminationCondition, EvidenceCalculation and
TerminationRegister of any jaxns or jaxns_amd module map to this package's types, and any other
NamedTuple is rebuilt as a plain collections.namedtuple with the stored fields.
"""
from __future__ import annotations

import base64
import collections
import io
import json
import warnings
from typing import Any, Dict, NamedTuple, Optional, Union

import numpy as np

from .types import (EvidenceCalculation, NestedSamplerResults, TerminationCondition, TerminationRegister)

_JAXNS_TYPES_MODULE = "jaxns.nested_samplers.common.types"
_KNOWN_TYPES = {t.__name__: t for t in (NestedSamplerResults, TerminationCondition, EvidenceCalculation,
                                        TerminationRegister)}


def _is_namedtuple(obj) -> bool:
    return isinstance(obj, tuple) and hasattr(obj, "_asdict") and hasattr(obj, "_fields")


def _class_path(obj) -> str:
    name = type(obj).__name__
    if type(obj) is _KNOWN_TYPES.get(name):
        return f"{_JAXNS_TYPES_MODULE}.{name}"
    return f"{type(obj).__module__}.{name}"


def _encode_array(a: np.ndarray) -> dict:
    a = np.array(a, copy=True, order="C")  # 0-d arrays stay 0-d
    return {"type": "__ndarray__", "__dtype__": str(a.dtype), "__data__": base64.b64encode(a.tobytes()).decode("ascii"),
            "__shape__": list(a.shape)}


def serialise_namedtuple(obj):
    """namedtuple_utils.py:34-48: the JSON-ready form of a NamedTuple of arrays, dicts and scalars."""
    if _is_namedtuple(obj):
        return {"type": "__namedtuple__", "__class__": _class_path(obj),
                "__data__": {k: serialise_namedtuple(v) for k, v in obj._asdict().items()}}
    if isinstance(obj, np.ndarray):
        return _encode_array(obj)
    if isinstance(obj, np.generic):
        return _encode_array(np.asarray(obj))
    if hasattr(obj, "__array__") and hasattr(obj, "dtype") and not isinstance(obj, (list, tuple, dict)):
        return _encode_array(np.asarray(obj))  # device tensors and the like
    if isinstance(obj, (list, tuple)):
        return [serialise_namedtuple(v) for v in obj]
    if isinstance(obj, dict):
        return {k: serialise_namedtuple(v) for k, v in obj.items()}
    return obj


def _decode_array(d: dict) -> np.ndarray:
    raw = base64.b64decode(d["__data__"])
    return np.frombuffer(raw, dtype=np.dtype(d["__dtype__"])).reshape(tuple(d["__shape__"])).copy()


def _namedtuple_type(path: str, fields):
    module, _, name = path.rpartition(".")
    if name in _KNOWN_TYPES and (module.startswith("jaxns.") or module.startswith("jaxns_amd.")):
        return _KNOWN_TYPES[name]
    return collections.namedtuple(name, list(fields))


def deserialise_namedtuple(obj):
    """namedtuple_utils.py:51-68, without importing the stored class path."""
    if isinstance(obj, dict):
        kind = obj.get("type")
        if kind == "__namedtuple__":
            data = obj["__data__"]
            cls = _namedtuple_type(obj["__class__"], data.keys())
            return cls(**{k: deserialise_namedtuple(v) for k, v in data.items()})
        if kind in ("__ndarray__", "__jax_ndarray__"):
            return _decode_array(obj)
        return {k: deserialise_namedtuple(v) for k, v in obj.items()}
    if isinstance(obj, list):
        return [deserialise_namedtuple(v) for v in obj]
    return obj


def save_pytree(pytree: NamedTuple, save_file: str):
    """utils.py:588-599."""
    if not _is_namedtuple(pytree):
        raise ValueError(f"Expected NamedTuple, got {type(pytree)}")
    with open(save_file, "w") as fp:
        json.dump(serialise_namedtuple(pytree), fp, indent=2)


def save_results(results: NestedSamplerResults, save_file: str):
    """utils.py:602-613."""
    if not save_file.lower().endswith(".json"):
        warnings.warn(f"Filename {save_file} does not end with .json. "
                      f"We let this pass, but you should consider using a .json file extension.")
    save_pytree(results, save_file)


def load_pytree(save_file: str):
    """utils.py:616-628."""
    with open(save_file, "r") as fp:
        return deserialise_namedtuple(json.load(fp))


def load_results(save_file: str) -> NestedSamplerResults:
    """utils.py:631-641."""
    return load_pytree(save_file)


# ----------------------------------------------------------------------------- summaries
def resample(key, samples, log_weights, S: Optional[int] = None, replace: bool = True):
    """utils.py:44-60 (resample_indicies, internals/random.py:34-84): S draws (rows) of `samples` (an
    array or a dict of arrays sharing the leading axis) with probabilities exp(log_weights), normalised;
    S defaults to Kish's ESS of the weights.  ``key`` is an integer seed."""
    lw = np.asarray(log_weights, dtype=np.float64)
    m = np.max(lw)
    p = np.exp(lw - m) if np.isfinite(m) else np.ones_like(lw)
    p = p / p.sum()
    if S is None:
        S = int(1.0 / np.sum(p * p))
    rng = np.random.default_rng(int(np.asarray(key).ravel()[-1]))
    idx = rng.choice(p.size, size=int(S), replace=replace, p=p)
    if isinstance(samples, dict):
        return {k: np.asarray(v)[idx] for k, v in samples.items()}
    return np.asarray(samples)[idx]


def maximum_a_posteriori_point(results: NestedSamplerResults) -> Dict[str, Any]:
    """utils.py:214-229: the samples at the largest log L + log prior."""
    i = int(np.argmax(results.log_posterior_density))
    return {k: np.asarray(v)[i] for k, v in results.samples.items()}


def maximum_a_posteriori_point_U(results: NestedSamplerResults) -> np.ndarray:
    """utils.py:232-246."""
    return np.asarray(results.U_samples)[int(np.argmax(results.log_posterior_density))]


_TERMINATION_TEXT = [
    "Reached max samples",
    "Evidence uncertainty low enough",
    "Small remaining evidence",
    "Reached ESS",
    "Used max num likelihood evaluations",
    "Likelihood contour reached",
    "Sampler efficiency too low",
    "All live-points are on a single plateau (sign of possible precision error)",
    "relative spread of live points < rtol",
    "absolute spread of live points < atol",
    "no seed points left (consider decreasing shell_fraction)",
    "XL < max(XL) * peak_XL_frac",
]


def _sig_round(v, uncert):
    """v rounded to two significant figures of `uncert` (utils.py:321-328)."""
    v, uncert = float(v), float(uncert)
    try:
        return round(v, -int("{:e}".format(uncert).split("e")[1]) + 1)
    except (ValueError, OverflowError):
        return v


def summary(results: NestedSamplerResults, with_parametrised: bool = False,
            f_obj: Optional[Union[str, io.TextIOBase]] = None) -> str:
    """utils.py:284-430: termination reasons, counts, log Z, max(log L), H, ESS, and per variable the
    mean, standard deviation and 10/50/90 percentiles of ESS-many posterior resamples (seed 23426),
    the MAP estimate and the max(L) estimate.  Printed when f_obj is None, else written to it; the
    text is also returned."""
    samples = dict(results.samples)
    if with_parametrised:
        samples.update(results.parametrised_samples)
    n = int(results.total_num_samples)
    ess = 100 if np.isnan(float(results.ESS)) else int(float(results.ESS))
    samples = {k: np.asarray(v)[:n] for k, v in samples.items()}
    log_L = np.asarray(results.log_L_samples)[:n]
    log_dp = np.asarray(results.log_dp_mean)[:n]
    i_max_L = int(np.argmax(log_L))
    i_map = int(np.argmax(np.asarray(results.log_posterior_density)[:n]))
    uniform = resample(23426, samples, log_dp, S=max(100, ess), replace=True)
    lines = []
    lines += ["--------", "Termination Conditions:"]
    reasons = np.atleast_1d(np.asarray(results.termination_reason))
    for si, reason in enumerate(reasons):
        if reasons.size > 1:
            lines.append(f"Sampler {si}:")
        for bit, text in enumerate(_TERMINATION_TEXT):
            if int(reason) >> bit & 1:
                lines.append(text)
    evals = int(results.total_num_likelihood_evaluations)
    phantom = int(results.total_phantom_samples)
    lu = float(results.log_Z_uncert)
    lines += [
        "--------",
        f"likelihood evals: {evals:d}",
        f"samples: {n:d}",
        f"phantom samples: {phantom:d}",
        f"likelihood evals / sample: {evals / n:.1f}",
        f"phantom fraction (%): {100.0 * phantom / n:.1f}%",
        "--------",
        f"logZ={_sig_round(results.log_Z_mean, lu)} +- {_sig_round(lu, lu)}",
        f"max(logL)={_sig_round(log_L[i_max_L], lu)}",
        f"H={_sig_round(results.H_mean, 0.1)}",
        f"ESS={int(np.asarray(results.ESS, np.float64)) if np.isfinite(float(results.ESS)) else results.ESS}",
    ]
    for name, draws in uniform.items():
        draws = draws.reshape((draws.shape[0], -1))
        ml = samples[name][i_max_L].reshape(-1)
        mp = samples[name][i_map].reshape(-1)
        nd = draws.shape[1]
        lines.append("--------")
        lines.append(f"{name if nd == 1 else name + '[#]'}: mean +- std.dev. | 10%ile / 50%ile / 90%ile | "
                     f"MAP est. | max(L) est.")
        for d in range(nd):
            sd = float(np.std(draws[:, d]))
            pct = np.percentile(draws[:, d], [10, 50, 90])
            r = lambda x: _sig_round(x, sd)  # noqa: E731
            label = name if nd == 1 else f"{name}[{d}]"
            lines.append(f"{label}: {r(np.mean(draws[:, d]))} +- {r(sd)} | {r(pct[0])} / {r(pct[1])} / {r(pct[2])} | "
                         f"{r(mp[d])} | {r(ml[d])}")
    lines.append("--------")
    out = "\n".join(lines)
    if f_obj is None:
        print(out)
    elif isinstance(f_obj, str):
        with open(f_obj, "w") as f:
            f.write(out)
    elif isinstance(f_obj, io.TextIOBase):
        f_obj.write(out)
    else:
        raise TypeError(f"Invalid f_obj: {type(f_obj)}")
    return out


def sample_evidence(key, num_live_points_per_sample, log_L_samples, S: int = 100) -> np.ndarray:
    """Samples of log Z by simulating the shrinkage, T_i ~ Beta(n_i, 1), X_i = X_{i-1} T_i,
    Z = sum (X_{i-1} - X_i) L_i (utils.py:433-480).  ``key`` is an integer seed."""
    n = np.asarray(num_live_points_per_sample, dtype=np.float64)
    L = np.asarray(log_L_samples, dtype=np.float64)
    rng = np.random.default_rng(int(np.asarray(key).ravel()[-1]))
    with np.errstate(divide="ignore", invalid="ignore"):
        log_T = np.log(rng.uniform(size=(S, n.size))) / n[None, :]
        log_X = np.cumsum(log_T, axis=1)
        log_X_prev = np.concatenate([np.zeros((S, 1)), log_X[:, :-1]], axis=1)
        # log(X_prev - X) = log X_prev + log(1 - T)
        log_dX = log_X_prev + np.log1p(-np.exp(log_T))
        terms = log_dX + L[None, :]
        m = np.max(terms, axis=1, keepdims=True)
        m = np.where(np.isfinite(m), m, 0.0)
        return (m[:, 0] + np.log(np.sum(np.exp(terms - m), axis=1)))
