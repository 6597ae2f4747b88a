"""The sampler plugin interface of jaxns (samplers/abc.py:18-115) on the device.

``UniDimSliceSampler`` and ``UniformSampler`` keep the reference's constructor arguments and the
``pre_process`` / ``get_sample`` / ``post_process`` / ``num_phantom`` protocol.  ``pre_process`` puts the
live points of an ``EphemeralState`` on the device as a one-run batch (sorted by log L, as the shell loop
keeps them); ``get_samples`` draws many samples above a contour in one launch of the fused sampling
kernel (hipns_shell_sample: one chain per sample, seeds chosen among the live points above the contour);
``get_sample`` is the one-sample case.  The random stream is the kernel's Philox stream: ``key`` is an
integer seed or a pair (seed, shell index), and a sampler called with the seed, shell index, contour and
live points of a shell of a run reproduces that shell's samples.
"""
from __future__ import annotations

import ctypes as C
from typing import Any, NamedTuple, Optional, Tuple

import numpy as np

from . import _lib
from ._lib import RunState, check, lib, ptr
from .engine import slice_sampler_code
from .types import TerminationRegister


class Sample(NamedTuple):
    """nested_samplers/common/types.py:101-105."""
    U_sample: np.ndarray
    log_L_constraint: np.ndarray
    log_L: np.ndarray
    num_likelihood_evaluations: np.ndarray


class LivePointCollection(NamedTuple):
    """nested_samplers/common/types.py:108-113."""
    sender_node_idx: np.ndarray
    U_sample: np.ndarray
    log_L_constraint: np.ndarray
    log_L: np.ndarray
    num_likelihood_evaluations: np.ndarray


class EphemeralState(NamedTuple):
    """samplers/abc.py:10-16."""
    key: Any
    live_points_collection: LivePointCollection
    termination_register: Optional[TerminationRegister] = None


class AbstractSampler:
    """samplers/abc.py:18-115, with a batched get_samples beside get_sample."""

    def pre_process(self, ephemeral_state: EphemeralState) -> Any:
        return self._pre_process(ephemeral_state)

    def post_process(self, ephemeral_state: EphemeralState, sampler_state: Any) -> Any:
        return self._post_process(ephemeral_state, sampler_state)

    def get_sample(self, key, log_L_constraint, sampler_state) -> Tuple[Sample, Sample]:
        s, ph = self.get_samples(key, log_L_constraint, sampler_state, 1)
        return Sample(*(np.asarray(x)[0] for x in s)), ph

    def get_samples(self, key, log_L_constraint, sampler_state, num_samples: int) -> Tuple[Sample, Sample]:
        raise NotImplementedError

    def num_phantom(self) -> int:
        raise NotImplementedError

    def _pre_process(self, ephemeral_state):
        raise NotImplementedError

    def _post_process(self, ephemeral_state, sampler_state):
        return sampler_state


class _DeviceSamplerBase(AbstractSampler):
    _sampler_code = 0

    def __init__(self, model, num_phantom_save: int = 0, midpoint_shrink: bool = True, num_slices: int = 1,
                 max_likelihood_evals: int = 100, lanes_per_chain: int = 2, waves_per_block: int = 2,
                 max_steps: int = 512):
        self.model = model
        self.num_slices = int(num_slices)
        self.num_phantom_save = int(num_phantom_save)
        self.midpoint_shrink = bool(midpoint_shrink)
        self.max_likelihood_evals = int(max_likelihood_evals)
        self._launch_kw = dict(lanes_per_chain=lanes_per_chain, waves_per_block=waves_per_block,
                               max_steps=max_steps)

    def _pre_process(self, ephemeral_state: EphemeralState):
        """The live points as a sorted one-run device batch (capacity grows on demand in get_samples)."""
        lp = ephemeral_state.live_points_collection
        U = np.atleast_2d(np.asarray(lp.U_sample, dtype=np.float64))
        L = np.asarray(lp.log_L, dtype=np.float64).ravel()
        order = np.argsort(L, kind="stable")
        return dict(U=U[order], log_L=L[order], batch=None, cap=0)

    def _batch_for(self, state, n: int):
        """A one-run batch holding the live points (padded below with -inf points, never seeds) for n
        chains: N = max(live points, n), shell = n."""
        import torch
        from .engine import DeviceBatch, RunSpec
        N_live = state["log_L"].size
        N = max(N_live, n, 2)
        if state["batch"] is not None and state["cap"] == (N, n):
            return state["batch"]
        k = self.num_phantom()
        spec = RunSpec(model=self.model, num_live_points=N, num_slices=self.num_slices, k_phantom=k,
                       max_samples=n * (1 + k) + N, seed=0, midpoint_shrink=self.midpoint_shrink,
                       shell_fraction=n / N)
        b = DeviceBatch([spec], **self._launch_kw)
        d = b.runs_host[0]
        d.shell = n
        d.sampler = self._sampler_code
        d.max_uniform_evals = self.max_likelihood_evals
        b.runs_dev.copy_(torch.frombuffer(bytearray(bytes(b.runs_host)), dtype=torch.uint8))
        pad = N - N_live
        L = np.concatenate([np.full(pad, -np.inf), state["log_L"]])
        U = np.full((N, b.dps), 0.5, np.float32)
        U[pad:, :self.model.U_ndims] = state["U"]
        b.live_U[0][:N] = torch.from_numpy(U)
        b.live_logL[0][:N] = torch.from_numpy(L)
        state["batch"], state["cap"] = b, (N, n)
        return b

    def get_samples(self, key, log_L_constraint, sampler_state, num_samples: int) -> Tuple[Sample, Sample]:
        import torch
        from .engine import seed_words
        n = int(num_samples)
        if n < 1:
            raise ValueError("num_samples must be >= 1")
        b = self._batch_for(sampler_state, n)
        d = b.runs_host[0]
        seed, iteration = (key if isinstance(key, tuple) else (key, 0))
        d.seed_lo, d.seed_hi = seed_words(seed)
        b.runs_dev.copy_(torch.frombuffer(bytearray(bytes(b.runs_host)), dtype=torch.uint8))
        L = b.live_logL[0][:int(d.N)].cpu().numpy()
        contour = float(log_L_constraint)
        st = RunState()
        st.buf = 0
        st.iteration = int(iteration)
        st.contour = contour
        st.first_seed = int(np.searchsorted(L, contour, side="right"))
        st.sender_node = -1
        b.states_dev.copy_(torch.frombuffer(bytearray(bytes(st)), dtype=torch.uint8))
        check(lib().hipns_shell_sample(ptr(b.runs_dev), ptr(b.states_dev), 1, b.runs_host, C.byref(b.bufs),
                                       C.byref(b.launch), _lib.stream_ptr()), "hipns_shell_sample")
        D = self.model.U_ndims
        s = Sample(U_sample=b.new_U[:n, :D].double().cpu().numpy(), log_L_constraint=np.full(n, contour),
                   log_L=b.new_logL[:n].cpu().numpy(),
                   num_likelihood_evaluations=b.new_evals[:n].to(torch.int64).cpu().numpy())
        k = self.num_phantom()
        if k:
            o = n  # the phantoms follow the shell in the collection (chain-major, k per chain)
            ph = Sample(U_sample=b.samp_U[o:o + n * k, :D].double().cpu().numpy(),
                        log_L_constraint=np.full(n * k, contour), log_L=b.samp_logL[o:o + n * k].cpu().numpy(),
                        num_likelihood_evaluations=np.zeros(n * k, np.int64))
        else:
            ph = Sample(np.zeros((0, D)), np.zeros(0), np.zeros(0), np.zeros(0, np.int64))
        return s, ph


class UniDimSliceSampler(_DeviceSamplerBase):
    """samplers/uni_slice_sampler.py:276-441: num_slices slices per sample along random directions (or,
    gradient_guided, each direction after the first reflected about d logL / dU at the last accepted
    point, or, gradient_slice, each slice along the unit gradient at its seed over [0, right]), the
    bracket the chord of the unit cube (perfect), midpoint (alpha) shrinkage, the last num_phantom_save
    intermediate points kept as phantom samples."""
    _sampler_code = 0

    def __init__(self, model, num_slices: int, num_phantom_save: int, midpoint_shrink: bool, perfect: bool = True,
                 gradient_slice: bool = False, gradient_guided: bool = False, **launch_kw):
        if not perfect:
            raise NotImplementedError("Step out not implemented.")  # as the reference (uni_slice_sampler.py:224)
        if (gradient_guided or gradient_slice) and \
                getattr(model.log_likelihood, "family", None) not in (_lib.FAM_MVN, _lib.FAM_GMM):
            raise NotImplementedError("gradient-guided or gradient-slice samples in one launch need the MVN or "
                                      "mixture family; other likelihoods run gradient-guided through "
                                      "NestedSampler(gradient_guided=True)")
        if num_slices < 1:
            raise ValueError(f"num_slices should be >= 1, got {num_slices}.")
        if not (0 <= num_phantom_save < num_slices):
            raise ValueError(f"num_phantom_save should be in [0, num_slices), got {num_phantom_save}.")
        super().__init__(model, num_phantom_save=num_phantom_save, midpoint_shrink=midpoint_shrink,
                         num_slices=num_slices, **launch_kw)
        self.gradient_guided = bool(gradient_guided)
        self.gradient_slice = bool(gradient_slice)
        # gradient_guided (code 2): after each accepted slice, d logL / dU at the accepted point (one more
        # evaluation) and the next direction reflected about it (uni_slice_sampler.py:255-271);
        # gradient_slice (code 3, 4 with both): each slice along the unit gradient at its seed (one more
        # evaluation) over [0, right] (uni_slice_sampler.py:202-214)
        self._sampler_code = slice_sampler_code(self.gradient_guided, self.gradient_slice)

    def num_phantom(self) -> int:
        return self.num_phantom_save


class UniformSampler(_DeviceSamplerBase):
    """samplers/uniform_samplers.py:20-85: rejection from the prior until log L >= the contour, at most
    max_likelihood_evals draws."""
    _sampler_code = 1

    def __init__(self, model, max_likelihood_evals: int = 100, **launch_kw):
        if max_likelihood_evals < 1:
            raise ValueError("max_likelihood_evals must be >= 1")
        super().__init__(model, max_likelihood_evals=max_likelihood_evals, num_slices=1, **launch_kw)

    def num_phantom(self) -> int:
        return 0
