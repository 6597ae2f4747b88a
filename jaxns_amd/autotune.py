"""Launch-shape autotuner for the fused sampling kernel.

Candidates are (lanes_per_chain, waves_per_block): 2 lanes per chain = 32
chains per wavefront with 32x32x2 FP32 MFMA tiles, 4 lanes per chain = 16
chains per wavefront with 16x16x4 tiles; 1-4 wavefronts per workgroup.  Each
candidate runs a few shells of a trial batch of the real shape (same model,
num_live_points, slices and number of runs) and is scored by likelihood
evaluations per second of the sampling kernel, timed with HIP events on the
launch stream.  Results are cached per (family, D, N, slices, runs, device).
"""
from __future__ import annotations

import ctypes as C
import json
import os
from typing import Iterable, Optional, Tuple

import numpy as np

CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), ".autotune_cache.json")
DEFAULT_CANDIDATES = ((2, 1), (2, 2), (2, 4), (4, 1), (4, 2), (4, 4))


def _key(model, N, num_slices, R):
    import torch
    name = torch.cuda.get_device_name(0) if torch.cuda.is_available() else "none"
    return f"fam{model.log_likelihood.family}-D{model.U_ndims}-N{N}-S{num_slices}-R{R}-{name}"


def measure(model, N, num_slices, R, lpc, wpb, shells=3, seed0=777) -> float:
    """Likelihood evaluations per second of the sampling kernel over `shells` shells."""
    import torch
    from ._lib import check, lib, make_term_cond, ptr, stream_ptr
    from .engine import DeviceBatch, RunSpec
    specs = [RunSpec(model=model, num_live_points=N, num_slices=num_slices, max_samples=N * (shells + 4),
                     seed=seed0 + r) for r in range(R)]
    b = DeviceBatch(specs, lanes_per_chain=lpc, waves_per_block=wpb)
    b.init_state()
    tc = make_term_cond(max_samples=float(N * (shells + 4)))
    s = stream_ptr()
    t = 0.0
    evals = 0
    prev = sum(st.num_likelihood_evaluations for st in b.states())
    for _ in range(shells):
        check(lib().hipns_shell_discard(ptr(b.runs_dev), ptr(b.states_dev), b.R, b.runs_host, C.byref(b.bufs), s))
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        check(lib().hipns_shell_sample(ptr(b.runs_dev), ptr(b.states_dev), b.R, b.runs_host, C.byref(b.bufs),
                                       C.byref(b.launch), s))
        e1.record()
        check(lib().hipns_shell_replace(ptr(b.runs_dev), ptr(b.states_dev), b.R, b.runs_host, C.byref(b.bufs), s))
        check(lib().hipns_shell_register(ptr(b.runs_dev), ptr(b.states_dev), b.R, b.runs_host, C.byref(b.bufs),
                                         C.byref(tc), s))
        torch.cuda.synchronize()
        t += e0.elapsed_time(e1) * 1e-3
        cur = sum(st.num_likelihood_evaluations for st in b.states())
        evals += cur - prev
        prev = cur
    return evals / t


def autotune(model, N: int, num_slices: int, R: int, candidates: Optional[Iterable[Tuple[int, int]]] = None,
             shells: int = 3, use_cache: bool = True, verbose: bool = False) -> Tuple[int, int]:
    key = _key(model, N, num_slices, R)
    cache = {}
    if use_cache and os.path.exists(CACHE):
        try:
            with open(CACHE) as f:
                cache = json.load(f)
        except (OSError, ValueError):
            cache = {}
    if key in cache:
        return tuple(cache[key]["best"])
    scores = {}
    for lpc, wpb in (candidates or DEFAULT_CANDIDATES):
        scores[(lpc, wpb)] = measure(model, N, num_slices, R, lpc, wpb, shells=shells)
        if verbose:
            print(f"[autotune] lpc={lpc} wpb={wpb}: {scores[(lpc, wpb)] / 1e9:.3f} G evals/s", flush=True)
    best = max(scores, key=scores.get)
    cache[key] = {"best": list(best), "scores": {f"{k[0]}x{k[1]}": v for k, v in scores.items()}}
    if use_cache:
        try:
            with open(CACHE, "w") as f:
                json.dump(cache, f, indent=1)
        except OSError:
            pass
    return best
