"""Launch-shape autotuner for the fused sampling kernel.

Candidates are (lanes_per_chain, waves_per_block): 2 lanes per chain = 32
chains per wavefront with 32x32 MFMA tiles, 4 lanes per chain = 16 chains per
wavefront with 16x16 tiles; the mixture family (no MFMA) also tries 8 lanes per
chain, 8 chains per wavefront with half the register slots of 4 lanes.  For the dense-Gaussian and mixture families the
workgroup holds 1-4 independent wavefronts of chains; for the regression family
its 1-4 wavefronts share one group of chains and split the data rows.  Each
candidate runs a few shells of a trial batch of the real shape (the same run
specs, with a small sample budget) and is scored by likelihood evaluations per
second of the sampling kernel, timed with HIP events on the launch stream.
For a batch of uneven runs (different dimensions or live-point counts) the trial's first shells
keep every run busy, while a whole step ends in sparse shells of the longest runs, which spread
better over more, smaller workgroups: there the fewest waves per workgroup among the candidates
within 5 % of the best score wins.  Results are cached per batch shape and device.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import hashlib
import json
import os
from typing import Iterable, List, Optional, Tuple

CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), ".autotune_cache.json")
DEFAULT_CANDIDATES = ((2, 1), (2, 2), (2, 4), (4, 1), (4, 2), (4, 4))
MIXTURE_CANDIDATES = DEFAULT_CANDIDATES + ((8, 1), (8, 2), (8, 4))


def _key(specs, lr_terms) -> str:
    import torch
    name = torch.cuda.get_device_name(0) if torch.cuda.is_available() else "none"
    lik = specs[0].model.log_likelihood
    fam = lik.family
    shape = [(s.model.U_ndims, s.num_live_points, s.num_slices) for s in specs]
    extra = ""
    if fam == 2:
        extra = f"-rows{lik.N}-t{lr_terms}"
    h = hashlib.sha1(repr(shape).encode()).hexdigest()[:12]
    return f"fam{fam}{extra}-R{len(specs)}-{h}-{name}-c4"


def measure(specs, lpc, wpb, shells=3, lr_terms=3, min_time=0.02, max_shells=12) -> float:
    """Likelihood evaluations per second of the sampling kernel over `shells` shells, or more (up to
    max_shells) until the timed kernel time reaches min_time seconds: short shells time noisily."""
    import torch
    from ._lib import check, lib, make_term_cond, ptr, stream_ptr
    from .engine import DeviceBatch
    trial = [dataclasses.replace(s, max_samples=s.num_live_points * (max(shells, max_shells) + 4), seed=777 + i)
             for i, s in enumerate(specs)]
    b = DeviceBatch(trial, lanes_per_chain=lpc, waves_per_block=wpb, lr_terms=lr_terms)
    b.init_state()
    tc = make_term_cond(max_samples=float(max(s.max_samples for s in trial)))
    s = stream_ptr()
    t = 0.0
    evals = 0
    prev = sum(st.num_likelihood_evaluations for st in b.states())
    done = 0
    while done < shells or (t < min_time and done < max_shells):
        done += 1
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


def autotune(specs: List, candidates: Optional[Iterable[Tuple[int, int]]] = None, shells: int = 3,
             use_cache: bool = True, verbose: bool = False, lr_terms: int = 3) -> Tuple[int, int]:
    """Best (lanes_per_chain, waves_per_block) for a batch of these run specs."""
    key = _key(specs, lr_terms)
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
    if candidates is None:
        from ._lib import FAM_GMM
        mix = specs[0].model.log_likelihood.family == FAM_GMM and max(s.model.U_ndims for s in specs) <= 128
        candidates = MIXTURE_CANDIDATES if mix else DEFAULT_CANDIDATES
    candidates = list(candidates)
    # one untimed trial first: the process's first shells pay one-time costs (code loading, first touch
    # of memory) that would otherwise count against the first candidate
    measure(specs, *candidates[0], shells=1, lr_terms=lr_terms)
    for lpc, wpb in candidates:
        scores[(lpc, wpb)] = measure(specs, lpc, wpb, shells=shells, lr_terms=lr_terms)
        if verbose:
            print(f"[autotune] lpc={lpc} wpb={wpb}: {scores[(lpc, wpb)] / 1e9:.3f} G evals/s", flush=True)
    best = max(scores, key=scores.get)
    if len({(s.model.U_ndims, s.num_live_points) for s in specs}) > 1:
        near = [c for c in scores if scores[c] >= 0.95 * scores[best]]
        best = min(near, key=lambda c: (c[1], -scores[c]))
    cache[key] = {"best": list(best), "scores": {f"{k[0]}x{k[1]}": v for k, v in scores.items()}}
    if use_cache:
        try:
            with open(CACHE, "w") as f:
                json.dump(cache, f, indent=1)
        except OSError:
            pass
    return best
