"""Post-run utilities with jaxns' names: summary, save_results, load_results (utils.py:284-641).

Results are stored as JSON with one key per NestedSamplerResults field
(arrays as nested lists, non-finite floats as strings), so that files written
here round-trip exactly.
"""
from __future__ import annotations

import io
import json
from typing import Optional, Union

import numpy as np

from .types import TERMINATION_BITS, NestedSamplerResults


def _enc(v):
    if isinstance(v, dict):
        return {"__dict__": {k: _enc(x) for k, x in v.items()}}
    if isinstance(v, np.ndarray):
        return {"__ndarray__": _enc(v.tolist()), "dtype": str(v.dtype), "shape": list(v.shape)}
    if isinstance(v, (list, tuple)):
        return [_enc(x) for x in v]
    if isinstance(v, (np.floating, float)):
        f = float(v)
        return f if np.isfinite(f) else {"__float__": repr(f)}
    if isinstance(v, (np.integer, int)):
        return int(v)
    return v


def _dec(v):
    if isinstance(v, dict):
        if "__dict__" in v:
            return {k: _dec(x) for k, x in v["__dict__"].items()}
        if "__ndarray__" in v:
            return np.array(_dec(v["__ndarray__"]), dtype=v["dtype"]).reshape(v["shape"])
        if "__float__" in v:
            return float(v["__float__"])
    if isinstance(v, list):
        return [_dec(x) for x in v]
    return v


def save_results(results: NestedSamplerResults, save_file: str):
    """utils.py:602-628."""
    with open(save_file, "w") as f:
        json.dump({k: _enc(getattr(results, k)) for k in results._fields}, f)


def load_results(save_file: str) -> NestedSamplerResults:
    """utils.py:631-641."""
    with open(save_file, "r") as f:
        d = json.load(f)
    return NestedSamplerResults(**{k: _dec(d[k]) for k in NestedSamplerResults._fields})


def summary(results: NestedSamplerResults, with_parametrised: bool = False, f_obj: Optional[Union[str, io.TextIOBase]] = None):
    """Text report of a run (utils.py:284-430)."""
    lines = ["--------", "Termination Conditions:"]
    reason = int(results.termination_reason)
    for bit, text in TERMINATION_BITS.items():
        if reason & (1 << bit):
            lines.append(text)
    lines += [
        "--------",
        f"likelihood evals: {int(results.total_num_likelihood_evaluations):d}",
        f"samples: {int(results.total_num_samples):d}",
        f"phantom samples: {int(results.total_phantom_samples):d}",
        f"likelihood evals / sample: {results.total_num_likelihood_evaluations / results.total_num_samples:.1f}",
        f"phantom fraction (%): {100.0 * results.total_phantom_samples / results.total_num_samples:.1f}%",
        "--------",
        f"logZ={results.log_Z_mean:.2f} +- {results.log_Z_uncert:.2f}",
        f"H={results.H_mean:.2f}",
        f"ESS={results.ESS:.0f}",
        "--------",
    ]
    log_w = np.asarray(results.log_dp_mean)
    w = np.exp(log_w - np.max(log_w)) if log_w.size else log_w
    w = w / np.sum(w) if w.size else w
    for name, x in results.samples.items():
        x = np.asarray(x)
        if x.ndim == 1:
            x = x[:, None]
        for i in range(x.shape[1]):
            m = float(np.sum(w * x[:, i]))
            sd = float(np.sqrt(max(np.sum(w * (x[:, i] - m) ** 2), 0.0)))
            lines.append(f"{name}[{i}]: mean +- std.dev. = {m:.2f} +- {sd:.2f}")
    lines.append("--------")
    out = "\n".join(lines)
    print(out)
    if f_obj is not None:
        if isinstance(f_obj, str):
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
