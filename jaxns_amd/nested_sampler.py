"""NestedSampler / DefaultNestedSampler with the jaxns public surface (public.py:31-222).

``ns = NestedSampler(model, ...)``; ``termination_reason, state = ns(key)``;
``results = ns.to_results(termination_reason, state)``.  ``key`` is an integer
seed (or a 2-word array); there is no JAX.  ``ns.run_batch(keys)`` advances
many seeds of the same model together on the device (one launch per kernel for
all runs, hipGraph-captured shells).
"""
from __future__ import annotations

import math
from typing import List, Optional, Sequence, Tuple

import numpy as np

from .engine import DeviceBatch, RunSpec, flatten_term_cond
from .framework import Model
from .types import NestedSamplerResults, TerminationCondition
from .utils import load_results, save_results, summary

INT64_MAX = float(np.iinfo(np.int64).max)


class NestedSamplerState:
    """A finished run held on the device, with the fields of NestedSamplerState (types.py:141-145):
    key, next_sample_idx, num_samples and sample_collection (copied to the host on access), plus the
    run's TerminationRegister."""

    def __init__(self, batch: DeviceBatch, run_index: int = 0):
        self.batch = batch
        self.run_index = run_index

    @property
    def key(self):
        return self.batch.specs[self.run_index].seed

    @property
    def next_sample_idx(self) -> int:
        return int(self.batch.states()[self.run_index].next_idx)

    @property
    def num_samples(self) -> int:
        return int(self.batch.states()[self.run_index].num_samples)

    @property
    def sample_collection(self):
        return self.batch.sample_collection(self.run_index)

    @property
    def termination_register(self):
        return self.batch.termination_register(self.run_index)


class NestedSampler:
    """public.py:31-222 -- static nested sampling with the uni-dimensional slice sampler.

    The reference's arguments keep their names, meanings and defaults.  Keyword-only extras:
    lanes_per_chain / waves_per_block / max_steps shape the sampler launch; lr_terms picks the
    regression family's product form (3 = BF16x3, 4 = BF16x3 plus the third weight part, 1 = plain
    FP32); refine_fp32 re-evaluates each shell's accepted points in plain FP32 before they enter the
    live set (regression family in a BF16 form)."""

    def __init__(self, model: Model, max_samples: Optional[float] = None, num_live_points: Optional[int] = None,
                 num_slices: Optional[int] = None, s: Optional[float] = None, k: Optional[int] = None,
                 c: Optional[int] = None, devices=None, difficult_model: bool = False,
                 parameter_estimation: bool = False, shell_fraction: float = 0.5, gradient_guided: bool = False,
                 init_efficiency_threshold: float = 0.1, verbose: bool = False, *, lanes_per_chain: int = 2,
                 waves_per_block: int = 4, max_steps: int = 512, lr_terms: int = 3, refine_fp32: bool = False):
        self.model = model
        D = model.U_ndims
        # public.py:69-107
        if num_slices is None:
            if difficult_model:
                s = 10 if s is None else float(s)
            else:
                s = 5 if s is None else float(s)
            if s <= 0:
                raise ValueError(f"Expected s > 0, got s={s}")
            num_slices = D * s
        self.s = s
        self.num_slices = int(num_slices)
        if parameter_estimation:
            max_k = self.s * D - 1
            k = min(D, max_k) if k is None else int(k)
        else:
            k = 0 if k is None else int(k)
        if not (0 <= k < self.num_slices):
            raise ValueError(f"Expected 0 <= k < num_slices, got k={k}, num_slices={self.num_slices}, U_ndims={D}")
        self.k = int(k)
        if num_live_points is not None:
            c = max(1, int(num_live_points / (self.k + 1)))
        else:
            if difficult_model:
                c = 100 * D if c is None else int(c)
            else:
                c = 30 * D if c is None else int(c)
            if c <= 0:
                raise ValueError(f"Expected c > 0, got c={c}")
        self.c = c
        if max_samples is None:
            max_samples = self.c * (self.k + 1) * 100
        self.max_samples = int(max_samples)
        self.gradient_guided = bool(gradient_guided)
        if devices is not None and len(devices) > 1:
            raise NotImplementedError("one GPU per process: shard independent runs across processes instead")
        self.difficult_model = difficult_model
        self.parameter_estimation = parameter_estimation
        self.shell_fraction = shell_fraction
        self.init_efficiency_threshold = init_efficiency_threshold
        self.verbose = verbose
        self.launch_kw = dict(lanes_per_chain=lanes_per_chain, waves_per_block=waves_per_block, max_steps=max_steps,
                              lr_terms=lr_terms, gradient_guided=self.gradient_guided, refine_fp32=refine_fp32)
        probe = self._spec(0)
        self.num_live_points = probe.num_live_points
        self._max_samples_rounded = probe.max_samples
        self.summary = summary
        self.save_results = save_results
        self.load_results = load_results

    def _spec(self, key) -> RunSpec:
        # ShardedStaticNestedSampler with num_live_points = c (public.py:109-126)
        return RunSpec(model=self.model, num_live_points=self.c, num_slices=self.num_slices, k_phantom=self.k,
                       max_samples=self.max_samples, seed=key, midpoint_shrink=not self.difficult_model,
                       shell_fraction=self.shell_fraction)

    def _term(self, term_cond: Optional[TerminationCondition]) -> dict:
        # public.py:153-170
        if term_cond is None:
            if self.parameter_estimation:
                term_cond = TerminationCondition(peak_XL_frac=0.1, max_samples=INT64_MAX)
            else:
                term_cond = TerminationCondition(dlogZ=math.log(1.0 + 1e-3), max_samples=INT64_MAX)
        ms = self._max_samples_rounded
        term_cond = term_cond._replace(
            max_samples=min(term_cond.max_samples, ms) if term_cond.max_samples is not None else ms)
        return flatten_term_cond(term_cond)

    def __call__(self, key, term_cond: Optional[TerminationCondition] = None) -> Tuple[int, NestedSamplerState]:
        batch = DeviceBatch([self._spec(key)], **self.launch_kw)
        batch.run(self._term(term_cond), init_efficiency_threshold=self.init_efficiency_threshold,
                  verbose=self.verbose)
        st = batch.states()[0]
        return int(st.termination_reason), NestedSamplerState(batch, 0)

    def run_batch(self, keys: Sequence, term_cond: Optional[TerminationCondition] = None, graph_iters: int = 8
                  ) -> Tuple[List[int], List[NestedSamplerState]]:
        """Advance len(keys) independent runs of this model together."""
        batch = DeviceBatch([self._spec(k) for k in keys], **self.launch_kw)
        batch.run(self._term(term_cond), init_efficiency_threshold=self.init_efficiency_threshold,
                  graph_iters=graph_iters)
        sts = batch.states()
        return [int(s.termination_reason) for s in sts], [NestedSamplerState(batch, r) for r in range(len(keys))]

    def to_results(self, termination_reason: int, state: NestedSamplerState, trim: bool = True
                   ) -> NestedSamplerResults:
        """public.py:177-196: trim=False keeps every row of the sample collection (max_samples), the
        statistics stopping at total_num_samples; trim_results cuts such results to that length."""
        return state.batch.to_results(state.run_index, trim=trim)

    @staticmethod
    def trim_results(results: NestedSamplerResults) -> NestedSamplerResults:
        """public.py:198-219: every per-sample array (size > 1) cut to total_num_samples."""
        n = int(results.total_num_samples)

        def trim(x):
            if isinstance(x, dict):
                return {k: trim(v) for k, v in x.items()}
            x = np.asarray(x)
            return x[:n] if x.size > 1 else x

        return NestedSamplerResults(**{k: trim(v) for k, v in results._asdict().items()})


DefaultNestedSampler = NestedSampler
