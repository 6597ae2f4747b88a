"""Device engine: a batch of nested-sampling runs advanced by libhipns kernels.

Layout in HBM (one allocation per field, runs packed back to back):
  live set   (double-buffered)  U [sum N, dps] f32 | logL, constraint f64 | sender, evals i32
  samples    U [sum max_samples, dps] f32 | logL f64 | sender, evals i32 | phantom u8
  new chains U [sum N, dps] f32 | logL f64 | evals i32
  params     packed float32 per-run model parameters
  runs       hipns_run_desc_t [R] (read-only)      states hipns_run_state_t [R]
``dps`` is the padded U-space stride (a multiple of 32).  Torch only provides
the device memory and the stream; every computation is a libhipns kernel.

The host loop mirrors ShardedStaticNestedSampler._run
(nested_samplers/sharded/sharded_static.py:775-851): create_init_state, the
uniform phase (whose termination check runs first and stops it before any
shell when efficiency_threshold > 0, as in the reference), the slice phase
until determine_termination, then the final live points are consumed.
"""
from __future__ import annotations

import ctypes as C
import math
import warnings
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _lib
from ._lib import Buffers, Launch, RunDesc, RunState, check, lib, ptr
from .types import (EvidenceCalculation, NestedSamplerResults, SampleCollection, TerminationCondition,
                    TerminationConditionConjunction, TerminationConditionDisjunction, TerminationRegister)

INT64_MAX = float(np.iinfo(np.int64).max)


def _torch():
    import torch
    return torch


def round_up_num_live_points(init_num_live_points, shell_frac, num_devices):
    """sharded_static.py:577-584."""
    n = int(init_num_live_points)
    while True:
        shell = int(n * shell_frac)
        if shell % num_devices == 0:
            return n
        n += 1


def round_up_max_samples(init_max_samples, num_discard, num_phantom_points):
    """sharded_static.py:587-594."""
    m = int(init_max_samples)
    block = num_discard * (1 + num_phantom_points)
    while m % block != 0:
        m += 1
    return m


@dataclass
class RunSpec:
    model: object
    num_live_points: int
    num_slices: int
    k_phantom: int = 0
    max_samples: int = 0
    seed: int = 0
    midpoint_shrink: bool = True
    shell_fraction: float = 0.5

    def __post_init__(self):
        self.shell_fraction = max(self.shell_fraction, 1.0 / self.num_live_points)
        self.num_live_points = round_up_num_live_points(self.num_live_points, self.shell_fraction, 1)
        self.shell = int(self.num_live_points * self.shell_fraction)
        self.max_samples = round_up_max_samples(self.max_samples, self.shell, self.k_phantom)


def flatten_term_cond(tc) -> dict:
    if isinstance(tc, (TerminationConditionConjunction, TerminationConditionDisjunction)):
        raise NotImplementedError("compound termination conditions are evaluated per flat condition only")
    d = {}
    for k in _lib._TC_NAMES:
        v = getattr(tc, k)
        if v is not None:
            d[k] = float(v)
    return d


def seed_words(key) -> tuple:
    a = np.asarray(key).ravel()
    if a.size >= 2:
        return int(a[0]) & 0xFFFFFFFF, int(a[1]) & 0xFFFFFFFF
    v = int(a[0]) if a.size else 0
    return v & 0xFFFFFFFF, (v >> 32) & 0xFFFFFFFF


def slice_sampler_code(gradient_guided: bool, gradient_slice: bool) -> int:
    """The run descriptor's sampler code of a UniDimSliceSampler: 0 plain, 2 gradient_guided, 3
    gradient_slice, 4 both."""
    if gradient_slice:
        return 4 if gradient_guided else 3
    return 2 if gradient_guided else 0


class DeviceBatch:
    """R independent runs sharing one likelihood family, advanced together."""

    def __init__(self, specs: List[RunSpec], *, lanes_per_chain: int = 2, waves_per_block: int = 4,
                 max_steps: int = 512, device: str = "cuda", lr_terms: int = 3, gradient_guided: bool = False,
                 refine_fp32: bool = False, slot_capacity: Optional[dict] = None, gradient_slice: bool = False):
        """gradient_slice: every slice along the unit gradient at its seed over the half bracket [0, right]
        (UniDimSliceSampler(gradient_slice=True), uni_slice_sampler.py:202-214), MVN and mixture families.
        slot_capacity (keys N, max_samples, D, num_slices, k_phantom): every run's buffers are sized for
        that capacity, so a finished run's slot can take another run of at most that size (run_queue);
        the shell launches are shaped for the capacity and read each run's own sizes on the device."""
        torch = _torch()
        if not specs:
            raise ValueError("no runs")
        self.specs = specs
        self.R = len(specs)
        fam = specs[0].model.log_likelihood
        self.family = getattr(fam, "family", None)
        self.gradient_guided = bool(gradient_guided)
        self.gradient_slice = bool(gradient_slice)
        if self.gradient_slice and self.family not in (_lib.FAM_MVN, _lib.FAM_GMM):
            raise NotImplementedError("gradient_slice runs in the fused kernel, for the MVN and mixture families")
        # gradient-guided slicing needs d logL / dU at accepted points: the MVN and mixture kernels compute
        # it themselves (sampler code 2, or 4 with gradient_slice); the regression family takes it from
        # torch autograd through the stepwise path
        self.external = self.family is None or (self.gradient_guided and self.family == _lib.FAM_LINREG)
        self.slice_sampler = slice_sampler_code(self.gradient_guided and not self.external, self.gradient_slice)
        if self.external:
            # a torch callable: evaluated between one-step sampler launches (hipns_ext_step)
            fn = fam.torch_log_prob if self.family is not None else fam
            if not callable(fn):
                raise TypeError("log_likelihood must be a likelihood family or a callable")
            if any(s.model is not specs[0].model for s in specs):
                raise ValueError("runs of the stepwise path must share one Model")
            from .likelihoods import TorchLikelihood
            self.ext_like = TorchLikelihood(fn, specs[0].model)
            self.family = _lib.FAM_MVN  # launch-struct placeholder; no fused family is launched
        for s in specs:
            if not self.external and getattr(s.model.log_likelihood, "family", None) != self.family:
                raise ValueError("all runs of a batch must share the likelihood family")
            if not self.external and s.model.dynamic_prior:
                raise ValueError("a prior that depends on other prior values needs a callable likelihood: the "
                                 "fused kernels evaluate a fixed affine prior")
        self.K = getattr(fam, "K", 1) if self.family == _lib.FAM_GMM else 1
        if self.family == _lib.FAM_GMM and any(getattr(s.model.log_likelihood, "K", 0) != self.K for s in specs):
            raise ValueError("all mixture runs of a batch must have the same number of components")
        self.slot_capacity = dict(slot_capacity) if slot_capacity else None
        if self.slot_capacity and (self.external or refine_fp32):
            raise ValueError("slot capacity needs a fused likelihood family without refinement")
        maxD = max(s.model.U_ndims for s in specs)
        if self.slot_capacity:
            maxD = max(maxD, int(self.slot_capacity["D"]))
        if maxD > _lib.MAX_DIM:
            raise ValueError(f"U dimension {maxD} exceeds {_lib.MAX_DIM}")
        self.dps = int(math.ceil(maxD / 32) * 32)
        self.device = torch.device(device)
        dev = self.device
        dps = self.dps
        if self.slot_capacity:
            cap = self.slot_capacity
            if any(s.num_live_points > cap["N"] or s.max_samples > cap["max_samples"] for s in specs):
                raise ValueError("a run exceeds the slot capacity")
            Ns = [int(cap["N"])] * self.R
            MSs = [int(cap["max_samples"])] * self.R
        else:
            Ns = [s.num_live_points for s in specs]
            MSs = [s.max_samples for s in specs]
        sumN, sumMS = sum(Ns), sum(MSs)
        # parameters
        prms, consts = [], []
        for s in specs:
            if self.external:
                p, c = np.zeros(4, np.float32), 0.0
            else:
                p, c = s.model.log_likelihood.device_params(s.model, dps)
            prms.append(p)
            consts.append(c)
        poffs = np.cumsum([0] + [int(p.numel()) if torch.is_tensor(p) else int(p.size) for p in prms])
        lr_ndata = lr_panels = lr_off = 0
        if self.family == _lib.FAM_LINREG:
            data = fam.data
            if any(s.model.log_likelihood.data is not data for s in specs):
                raise ValueError("all regression runs of a batch must share one RegressionData")
            blk, lr_ndata, lr_panels = data.packed(dps)
            lr_off = int(-(-poffs[-1] // 64) * 64)  # 256-byte aligned start of the shared rows
            prms.append(np.zeros(lr_off - poffs[-1], np.float32))
            prms.append(blk)
        # parameter blocks formed on the host (numpy) or on the device (a tensor, e.g. the FP64-prepared
        # multivariate normal) are packed back to back
        self.params = torch.cat([p.to(device=dev, dtype=torch.float32).reshape(-1) if torch.is_tensor(p)
                                 else torch.from_numpy(np.ascontiguousarray(p, dtype=np.float32)).to(dev)
                                 for p in prms])
        # descriptors
        arr = (RunDesc * self.R)()
        lo = so = 0
        for r, s in enumerate(specs):
            d = arr[r]
            d.D = s.model.U_ndims
            d.N = s.num_live_points
            d.shell = s.shell
            d.num_slices = s.num_slices
            d.k_phantom = s.k_phantom
            d.midpoint_shrink = int(s.midpoint_shrink)
            d.sampler = self.slice_sampler
            d.max_uniform_evals = 100
            d.max_samples = s.max_samples
            d.live_off = lo
            d.samp_off = so
            d.chain_off = lo
            d.param_off = int(poffs[r])
            d.seed_lo, d.seed_hi = seed_words(s.seed)
            d.log_const = consts[r]
            lo += Ns[r]
            so += MSs[r]
        self.runs_host = arr
        self.param_block = [int(poffs[r + 1] - poffs[r]) for r in range(self.R)]
        # descriptors the shell launches are shaped by: the runs themselves, or the slot capacity
        self.runs_launch = arr
        if self.slot_capacity:
            cap = self.slot_capacity
            lau = (RunDesc * self.R).from_buffer_copy(arr)
            for r in range(self.R):
                d = lau[r]
                d.D, d.N, d.shell = int(cap["D"]), int(cap["N"]), int(cap["N"]) // 2
                d.num_slices, d.k_phantom = int(cap["num_slices"]), int(cap["k_phantom"])
                d.max_samples = int(cap["max_samples"])
            self.runs_launch = lau
        self.runs_dev = torch.frombuffer(bytearray(bytes(arr)), dtype=torch.uint8).to(dev)
        self.states_dev = torch.zeros(self.R * C.sizeof(RunState), dtype=torch.uint8, device=dev)
        f32, f64, i32 = torch.float32, torch.float64, torch.int32
        self.live_U = [torch.full((sumN, dps), 0.5, dtype=f32, device=dev) for _ in range(2)]
        self.live_logL = [torch.zeros(sumN, dtype=f64, device=dev) for _ in range(2)]
        self.live_cons = [torch.zeros(sumN, dtype=f64, device=dev) for _ in range(2)]
        self.live_sender = [torch.zeros(sumN, dtype=i32, device=dev) for _ in range(2)]
        self.live_evals = [torch.zeros(sumN, dtype=i32, device=dev) for _ in range(2)]
        self.samp_U = torch.empty((sumMS, dps), dtype=f32, device=dev)
        self.samp_logL = torch.empty(sumMS, dtype=f64, device=dev)
        self.samp_sender = torch.empty(sumMS, dtype=i32, device=dev)
        self.samp_evals = torch.empty(sumMS, dtype=i32, device=dev)
        self.samp_phantom = torch.empty(sumMS, dtype=torch.uint8, device=dev)
        self.new_U = torch.full((sumN, dps), 0.5, dtype=f32, device=dev)
        self.new_logL = torch.zeros(sumN, dtype=f64, device=dev)
        self.new_evals = torch.zeros(sumN, dtype=i32, device=dev)
        b = Buffers()
        for i in range(2):
            b.live_U[i] = ptr(self.live_U[i])
            b.live_logL[i] = ptr(self.live_logL[i])
            b.live_cons[i] = ptr(self.live_cons[i])
            b.live_sender[i] = ptr(self.live_sender[i])
            b.live_evals[i] = ptr(self.live_evals[i])
        b.samp_U, b.samp_logL, b.samp_sender = ptr(self.samp_U), ptr(self.samp_logL), ptr(self.samp_sender)
        b.samp_evals, b.samp_phantom = ptr(self.samp_evals), ptr(self.samp_phantom)
        b.new_U, b.new_logL, b.new_evals = ptr(self.new_U), ptr(self.new_logL), ptr(self.new_evals)
        b.params = ptr(self.params)
        b.dps = dps
        # dispatch order of the sampling launch: runs with the longest chains (most slices) first, so
        # that in a batch of uneven runs the slowest workgroups do not start last (stable: ties keep
        # the run order, and a batch of equal runs keeps the identity)
        cost = np.array([s.num_slices * max(1, s.model.U_ndims) for s in specs], np.int64)
        order = np.argsort(-cost, kind="stable").astype(np.int32)
        self.run_order = torch.from_numpy(order).to(dev)
        b.run_order = ptr(self.run_order)
        self.bufs = b
        # the same buffers for calls on a subset of the runs (one slot, a gathered set): the dispatch
        # order indexes the whole batch, so those calls run their runs in the given order
        self.bufs_subset = Buffers.from_buffer_copy(b)
        self.bufs_subset.run_order = None
        if lr_terms not in (1, 3, 4):
            raise ValueError("lr_terms must be 1 (plain FP32), 3 (BF16x3) or 4")
        self.launch = Launch(self.family, lanes_per_chain, waves_per_block, self.K, lr_ndata, max_steps, lr_panels,
                             lr_terms, lr_off)
        # FP32 refinement of accepted points (regression family in a BF16 form): the BF16 product decides
        # acceptance, the plain FP32 form re-evaluates each shell's accepted points before they enter
        # the live set, so the stored logL (what the evidence integrates) has FP32 error
        self.refine_fp32 = bool(refine_fp32)
        if self.refine_fp32:
            if self.external or self.family != _lib.FAM_LINREG or lr_terms == 1:
                raise ValueError("refine_fp32 applies to the regression family in a BF16 form (lr_terms 3 or 4)")
            self.launch_fp32 = Launch.from_buffer_copy(self.launch)
            self.launch_fp32.lr_terms = 5  # plain FP32 product, w - w0 from the FP64 transform
            # per run: refined accepted points whose FP32 logL is not above the shell's contour (the
            # BF16 value that decided acceptance was); they enter the live set below the contour and
            # are discarded again by the next shell
            self.refine_below = torch.zeros(self.R, dtype=torch.int64, device=dev)
        if self.external:
            i32 = torch.int32
            self.ext_U0 = torch.full((sumN, dps), 0.5, dtype=f32, device=dev)
            self.ext_nv = torch.zeros((sumN, dps), dtype=f32, device=dev)
            self.ext_prop_U = torch.full((sumN, dps), 0.5, dtype=f32, device=dev)
            self.ext_logL0 = torch.zeros(sumN, dtype=f64, device=dev)
            self.ext_prop_logL = torch.zeros(sumN, dtype=f64, device=dev)
            self.ext_left = torch.zeros(sumN, dtype=f32, device=dev)
            self.ext_right = torch.zeros(sumN, dtype=f32, device=dev)
            self.ext_t = torch.zeros(sumN, dtype=f32, device=dev)
            self.ext_s = torch.zeros(sumN, dtype=i32, device=dev)
            self.ext_k = torch.zeros(sumN, dtype=i32, device=dev)
            self.ext_evals = torch.zeros(sumN, dtype=i32, device=dev)
            self.ext_flags = torch.zeros(sumN, dtype=i32, device=dev)
            self.ext_grad = torch.zeros((sumN, dps), dtype=f32, device=dev)
            e = _lib.ExtState()
            for name in ("U0", "nv", "prop_U", "logL0", "prop_logL", "left", "right", "t", "s", "k", "evals", "flags",
                         "grad"):
                setattr(e, name, ptr(getattr(self, "ext_" + name)))
            self.ext = e
        self._graph = None
        self._ev_ws = None
        self._graph_iters = 0
        self._graph_term = None

    # ------------------------------------------------------------------ calls
    def _s(self, stream):
        return _lib.stream_ptr(stream)

    def init_state(self, stream=None):
        if self.external:
            s = self._s(stream)
            check(lib().hipns_init_prepare(ptr(self.runs_dev), ptr(self.states_dev), self.R, self.runs_host,
                                           C.byref(self.bufs), s), "hipns_init_prepare")
            self._ext_steps(2, stream)
            check(lib().hipns_init_finalize(ptr(self.runs_dev), ptr(self.states_dev), self.R, self.runs_host,
                                            C.byref(self.bufs), s), "hipns_init_finalize")
            return
        check(lib().hipns_create_init_state(ptr(self.runs_dev), ptr(self.states_dev), self.R, self.runs_host,
                                            C.byref(self.bufs), C.byref(self.launch), None, 0, self._s(stream)),
              "hipns_create_init_state")

    def _ext_steps(self, mode: int, stream=None):
        """Step every chain until none runs, evaluating each step's proposals with the callable
        (Model.forward over the batch)."""
        s = self._s(stream)
        self.ext_flags.zero_()
        w = 0
        while True:
            check(lib().hipns_ext_step(ptr(self.runs_dev), ptr(self.states_dev), self.R, self.runs_host,
                                       C.byref(self.bufs), C.byref(self.launch), C.byref(self.ext), mode, w, s),
                  "hipns_ext_step")
            if not bool((self.ext_flags & 1).any()):
                break
            if mode == 4 and bool((self.ext_flags & 8).any()):
                torch = _torch()
                with torch.enable_grad():
                    U = self.ext_prop_U.detach().clone().requires_grad_(True)
                    L = self.ext_like(U)
                    g, = torch.autograd.grad(torch.where(torch.isfinite(L), L, torch.zeros_like(L)).sum(), U)
                self.ext_prop_logL.copy_(L.detach())
                self.ext_grad.copy_(torch.nan_to_num(g, nan=0.0, posinf=0.0, neginf=0.0))
            else:
                self.ext_prop_logL.copy_(self.ext_like(self.ext_prop_U))
            w += 1

    def check_termination(self, term: dict, set_no_seed: bool = True, stream=None):
        tc = _lib.make_term_cond(**term)
        check(lib().hipns_check_termination(ptr(self.runs_dev), ptr(self.states_dev), self.R, C.byref(self.bufs),
                                            C.byref(tc), int(set_no_seed), self._s(stream)),
              "hipns_check_termination")

    def refine_new(self, stream=None):
        """Re-evaluate the accepted points of the current shell (every run's new samples) with the plain
        FP32 regression product on w - w0 formed in FP64 (lr_terms 5) and overwrite their logL
        (hipns_log_likelihood, one launch per run).  Phantom samples keep the BF16 value; evaluation
        counts do not include the refinement."""
        s = self._s(stream)
        sz = C.sizeof(RunDesc)
        base_U, base_L, base_runs = ptr(self.new_U), ptr(self.new_logL), ptr(self.runs_dev)
        torch = _torch()
        contour = self.states_dev.view(torch.float64).view(self.R, -1)[:, RunState.contour.offset // 8]
        done = self.states_dev.view(torch.int32).view(self.R, -1)[:, RunState.done.offset // 4]
        for r in range(self.R):
            d = self.runs_host[r]
            n, off = int(d.shell), int(d.chain_off)
            check(lib().hipns_log_likelihood(base_U + off * self.dps * 4, n, base_L + off * 8, base_runs + r * sz,
                                             C.pointer(self.runs_host[r]), C.byref(self.bufs_subset),
                                             C.byref(self.launch_fp32), s), "hipns_log_likelihood")
            below = (self.new_logL[off:off + n] <= contour[r]).sum() * (done[r] == 0)
            self.refine_below[r:r + 1] += below

    def collect_shell(self, term: dict, stream=None):
        tc = _lib.make_term_cond(**term)
        if self.refine_fp32:
            s = self._s(stream)
            args = (ptr(self.runs_dev), ptr(self.states_dev), self.R, self.runs_host, C.byref(self.bufs))
            check(lib().hipns_shell_discard(*args, s), "hipns_shell_discard")
            check(lib().hipns_shell_sample(*args, C.byref(self.launch), s), "hipns_shell_sample")
            self.refine_new(stream)
            check(lib().hipns_shell_replace(*args, s), "hipns_shell_replace")
            check(lib().hipns_shell_register(*args, C.byref(tc), s), "hipns_shell_register")
            return
        if self.external:
            s = self._s(stream)
            args = (ptr(self.runs_dev), ptr(self.states_dev), self.R, self.runs_host, C.byref(self.bufs))
            check(lib().hipns_shell_discard(*args, s), "hipns_shell_discard")
            mode = 1 if self.runs_host[0].sampler == 1 else (4 if self.gradient_guided else 0)
            self._ext_steps(mode, stream)
            check(lib().hipns_shell_replace(*args, s), "hipns_shell_replace")
            check(lib().hipns_shell_register(*args, C.byref(tc), s), "hipns_shell_register")
            return
        check(lib().hipns_collect_shell(ptr(self.runs_dev), ptr(self.states_dev), self.R, self.runs_launch,
                                        C.byref(self.bufs), C.byref(self.launch), C.byref(tc), None, 0,
                                        self._s(stream)), "hipns_collect_shell")

    def consume(self, stream=None):
        check(lib().hipns_consume_live(ptr(self.runs_dev), ptr(self.states_dev), self.R, self.runs_host,
                                       C.byref(self.bufs), self._s(stream)), "hipns_consume_live")

    def states(self) -> List[RunState]:
        raw = self.states_dev.cpu().numpy().tobytes()
        sz = C.sizeof(RunState)
        return [RunState.from_buffer_copy(raw[i * sz:(i + 1) * sz]) for i in range(self.R)]

    def done_flags(self) -> np.ndarray:
        torch = _torch()
        st = self.states_dev.view(self.R, C.sizeof(RunState))
        off = RunState.done.offset
        flags = st[:, off:off + 4].contiguous().view(torch.int32).cpu().numpy().ravel()
        return flags

    # ------------------------------------------------------------------ slots (run_queue)
    def _slot_args(self, r: int):
        sd, ss = C.sizeof(RunDesc), C.sizeof(RunState)
        return (ptr(self.runs_dev) + r * sd, ptr(self.states_dev) + r * ss, 1, C.pointer(self.runs_host[r]),
                C.byref(self.bufs_subset))

    def assign(self, r: int, spec: RunSpec):
        """Give slot r (whose run has finished) the run `spec`: its descriptor and model parameters.
        The buffers keep the slot's offsets; call init_slot next."""
        torch = _torch()
        if self.slot_capacity is None:
            raise ValueError("assign needs a batch built with slot_capacity")
        cap = self.slot_capacity
        if spec.num_live_points > cap["N"] or spec.max_samples > cap["max_samples"] or \
                spec.model.U_ndims > cap["D"] or spec.num_slices > cap["num_slices"] or spec.k_phantom > cap["k_phantom"]:
            raise ValueError("run exceeds the slot capacity")
        lik = spec.model.log_likelihood
        if getattr(lik, "family", None) != self.family or (self.family == _lib.FAM_GMM and lik.K != self.K):
            raise ValueError("run of another likelihood family")
        p, c = lik.device_params(spec.model, self.dps)
        p = p.reshape(-1) if torch.is_tensor(p) else torch.from_numpy(np.ascontiguousarray(p, dtype=np.float32))
        if p.numel() != self.param_block[r]:
            raise ValueError("parameter block of another size")
        d = self.runs_host[r]
        o = int(d.param_off)
        self.params[o:o + p.numel()].copy_(p.to(self.device, torch.float32))
        d.D = spec.model.U_ndims
        d.N = spec.num_live_points
        d.shell = spec.shell
        d.num_slices = spec.num_slices
        d.k_phantom = spec.k_phantom
        d.midpoint_shrink = int(spec.midpoint_shrink)
        d.sampler = self.slice_sampler
        d.max_samples = spec.max_samples
        d.seed_lo, d.seed_hi = seed_words(spec.seed)
        d.log_const = c
        sd = C.sizeof(RunDesc)
        self.runs_dev[r * sd:(r + 1) * sd].copy_(torch.frombuffer(bytearray(bytes(d)), dtype=torch.uint8))
        self.specs[r] = spec

    def init_slot(self, r: int, term: dict, init_efficiency_threshold: float = 0.1, stream=None):
        """create_init_state for slot r alone, then the termination checks that start a run
        (DeviceBatch.run): the uniform phase's, which stops before its first shell for any threshold > 0,
        then `term`'s."""
        s = self._s(stream)
        args = self._slot_args(r)
        check(lib().hipns_create_init_state(*args, C.byref(self.launch), None, 0, s), "hipns_create_init_state")
        if init_efficiency_threshold > 0:
            ut = _lib.make_term_cond(efficiency_threshold=float(init_efficiency_threshold), dlogZ=0.0,
                                     max_samples=float(self.runs_host[r].max_samples))
            check(lib().hipns_check_termination(args[0], args[1], 1, args[4], C.byref(ut), 1, s),
                  "hipns_check_termination")
        tc = _lib.make_term_cond(**term)
        check(lib().hipns_check_termination(args[0], args[1], 1, args[4], C.byref(tc), 1, s),
              "hipns_check_termination")

    def consume_slot(self, r: int, stream=None):
        check(lib().hipns_consume_live(*self._slot_args(r), self._s(stream)), "hipns_consume_live")

    def log_Z_slots(self, slots: List[int], stream=None) -> np.ndarray:
        """log_Z of the runs in the given slots (their descriptors and states gathered into a compact
        batch for hipns_batch_evidence)."""
        torch = _torch()
        sd, ss = C.sizeof(RunDesc), C.sizeof(RunState)
        idx = torch.as_tensor(slots, dtype=torch.int64, device=self.device)
        runs = self.runs_dev.view(self.R, sd).index_select(0, idx).contiguous()
        sts = self.states_dev.view(self.R, ss).index_select(0, idx).contiguous()
        host = (RunDesc * len(slots))(*[self.runs_host[r] for r in slots])
        if self._ev_ws is None:
            nbytes = int(lib().hipns_batch_evidence_workspace_size(self.R, self.runs_launch))
            self._ev_ws = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        n = len(slots)
        fins = torch.empty((n, 8), dtype=torch.float64, device=self.device)
        phs = torch.empty(n, dtype=torch.int64, device=self.device)
        check(lib().hipns_batch_evidence(ptr(runs), ptr(sts), n, host, C.byref(self.bufs_subset), ptr(fins),
                                         ptr(phs), ptr(self._ev_ws), self._ev_ws.numel(), self._s(stream)),
              "hipns_batch_evidence")
        fin, ph = fins.cpu().numpy(), phs.cpu().numpy()
        raw = sts.cpu().numpy().tobytes()
        out = []
        for i, r in enumerate(slots):
            st = RunState.from_buffer_copy(raw[i * ss:(i + 1) * ss])
            M = int(min(st.num_samples, self.runs_host[r].max_samples))
            mu, var = _linear_to_log_stats(fin[i, 3], fin[i, 5])
            pf = ph[i] / max(M, 1)
            k = pf / (1.0 - pf)
            out.append((mu, math.sqrt(var) * math.sqrt(1.0 + k)))
        return np.array(out)

    def termination_register(self, r: int) -> TerminationRegister:
        """TerminationRegister of run r (nested_samplers/common/types.py:123-134) from its device state."""
        st = self.states()[r]
        return TerminationRegister(
            num_samples_used=int(st.num_samples_used),
            evidence_calc=EvidenceCalculation(*[float(v) for v in st.evidence_calc]),
            evidence_calc_with_remaining=EvidenceCalculation(*[float(v) for v in st.evidence_calc_with_remaining]),
            num_likelihood_evaluations=int(st.num_likelihood_evaluations), log_L_contour=float(st.log_L_contour),
            efficiency=float(st.efficiency), plateau=bool(st.plateau), no_seed_points=bool(st.no_seed_points),
            relative_spread=float(st.relative_spread), absolute_spread=float(st.absolute_spread),
            peak_log_XL=float(st.peak_log_XL))

    def sample_collection(self, r: int) -> SampleCollection:
        """Run r's sample collection (all max_samples entries) as host arrays."""
        d = self.runs_host[r]
        o, ms = int(d.samp_off), int(d.max_samples)
        return SampleCollection(sender_node_idx=self.samp_sender[o:o + ms].cpu().numpy(),
                                log_L=self.samp_logL[o:o + ms].cpu().numpy(),
                                U_samples=self.samp_U[o:o + ms, :int(d.D)].cpu().numpy(),
                                num_likelihood_evaluations=self.samp_evals[o:o + ms].cpu().numpy(),
                                phantom=self.samp_phantom[o:o + ms].cpu().numpy().astype(bool))

    def report(self, r: int = 0) -> str:
        """The per-shell report of _main_ns_thread with verbose=True (sharded_static.py:519-549)."""
        st = self.states()[r]
        ewr, ec = st.evidence_calc_with_remaining, st.evidence_calc
        mu, var = _linear_to_log_stats(ewr[3], ewr[5])
        mu0, var0 = _linear_to_log_stats(ec[3], ec[5])
        ess = math.exp((ewr[3] + ewr[3]) - ewr[7]) if np.isfinite(ewr[7]) else float("nan")
        return ("-------\n"
                f"Num samples: {st.num_samples_used}\n"
                f"Num likelihood evals: {st.num_likelihood_evaluations}\n"
                f"Efficiency: {st.efficiency}\n"
                f"log(L) contour: {st.log_L_contour}\n"
                f"log(Z) est.: {mu} +- {math.sqrt(var)}\n"
                f"log(Z | remaining) est.: {mu - mu0} +- {math.sqrt(var + var0)}\n"
                f"ESS: {ess}\n")

    def cap_hits(self) -> np.ndarray:
        """Per run: slices that reached the step cap (max_steps shrinks) and kept their seed point.
        The reference's shrink loop has no cap, so a non-zero count means the run left its algorithm."""
        return np.array([int(st.cap_hits) for st in self.states()], dtype=np.int64)

    def set_sampler(self, sampler: int):
        for r in range(self.R):
            self.runs_host[r].sampler = sampler
        torch = _torch()
        self.runs_dev.copy_(torch.frombuffer(bytearray(bytes(self.runs_host)), dtype=torch.uint8))

    # ------------------------------------------------------------------ loop
    def run(self, term: dict, *, init_efficiency_threshold: float = 0.1, graph_iters: int = 0,
            max_iterations: int = 1_000_000, stream=None, verbose: bool = False):
        """create_init_state -> phases -> consume.  graph_iters > 0 replays a captured
        hipGraph of that many shells between host checks; 0 syncs after every shell.  verbose prints
        run 0's register after every host check (every shell without a graph)."""
        torch = _torch()
        self._verbose = verbose
        self.init_state(stream)
        if init_efficiency_threshold > 0:
            uterm = dict(efficiency_threshold=float(init_efficiency_threshold), dlogZ=0.0,
                         max_samples=float(min(s.max_samples for s in self.specs)))
            self.check_termination(uterm, True, stream)
            if not np.all(self.done_flags()):
                self.set_sampler(1)
                self._loop(uterm, graph_iters=0, max_iterations=max_iterations, stream=stream)
                self.set_sampler(self.slice_sampler)
        self.check_termination(term, True, stream)
        self._loop(term, graph_iters=graph_iters, max_iterations=max_iterations, stream=stream)
        self.consume(stream)
        torch.cuda.synchronize()

    def _loop(self, term, *, graph_iters, max_iterations, stream):
        torch = _torch()
        it = 0
        if graph_iters and graph_iters > 0 and not self.external:
            g = self._get_graph(term, graph_iters)
            while not np.all(self.done_flags()) and it < max_iterations:
                g.replay()
                it += graph_iters
                if getattr(self, "_verbose", False):
                    print(self.report(0), flush=True)
        else:
            while it < max_iterations:
                if np.all(self.done_flags()):
                    break
                self.collect_shell(term, stream)
                it += 1
                if getattr(self, "_verbose", False):
                    print(self.report(0), flush=True)

    def _get_graph(self, term, iters):
        torch = _torch()
        key = (tuple(sorted(term.items())), iters)
        if self._graph is not None and self._graph_term == key:
            return self._graph
        s = torch.cuda.Stream()
        g = torch.cuda.CUDAGraph()
        torch.cuda.synchronize()
        with torch.cuda.graph(g, stream=s, capture_error_mode="relaxed"):
            for _ in range(iters):
                self.collect_shell(term, stream=s)
        self._graph = g
        self._graph_term = key
        return g

    # ------------------------------------------------------------------ results
    def log_Z(self, stream=None) -> np.ndarray:
        """Per-run (log_Z_mean, log_Z_uncert) of _to_results (sharded_static.py:706-745): the tree
        count and evidence scan of every run in one set of launches (hipns_batch_evidence), one
        host copy at the end."""
        torch = _torch()
        dev = self.device
        if self._ev_ws is None:
            nbytes = int(lib().hipns_batch_evidence_workspace_size(self.R, self.runs_host))
            self._ev_ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        fins = torch.empty((self.R, 8), dtype=torch.float64, device=dev)
        phs = torch.empty(self.R, dtype=torch.int64, device=dev)
        check(lib().hipns_batch_evidence(ptr(self.runs_dev), ptr(self.states_dev), self.R, self.runs_host,
                                         C.byref(self.bufs), ptr(fins), ptr(phs), ptr(self._ev_ws),
                                         self._ev_ws.numel(), self._s(stream)), "hipns_batch_evidence")
        fin = fins.cpu().numpy()
        ph = phs.cpu().numpy()
        sts = self.states()
        out = []
        for r in range(self.R):
            M = int(min(sts[r].num_samples, self.runs_host[r].max_samples))
            mu, var = _linear_to_log_stats(fin[r, 3], fin[r, 5])
            pf = ph[r] / max(M, 1)
            k = pf / (1.0 - pf)
            out.append((mu, math.sqrt(var) * math.sqrt(1.0 + k)))
        return np.array(out)

    def log_Z_per_run(self, stream=None) -> np.ndarray:
        """log_Z through the single-run primitives (hipns_count_crossed_edges + hipns_evidence_scan)
        for every run in turn; the batched path must agree with it exactly."""
        torch = _torch()
        sts = self.states()
        dev = self.device
        Ms = [int(min(st.num_samples, self.runs_host[r].max_samples)) for r, st in enumerate(sts)]
        Mmax = max(Ms)
        ws_c = torch.empty(int(lib().hipns_count_crossed_edges_workspace_size(Mmax)), dtype=torch.uint8, device=dev)
        ws_e = torch.empty(int(lib().hipns_evidence_scan_workspace_size(Mmax)), dtype=torch.uint8, device=dev)
        si = torch.empty(Mmax, dtype=torch.int64, device=dev)
        nl = torch.empty(Mmax, dtype=torch.int64, device=dev)
        lL = torch.empty(Mmax, dtype=torch.float64, device=dev)
        nld = torch.empty(Mmax, dtype=torch.float64, device=dev)
        fins = torch.empty((self.R, 8), dtype=torch.float64, device=dev)
        phs = torch.zeros(self.R, dtype=torch.int64, device=dev)
        s = self._s(stream)
        for r in range(self.R):
            M = Ms[r]
            off = int(self.runs_host[r].samp_off)
            logL = self.samp_logL[off:off + M]
            check(lib().hipns_count_crossed_edges(ptr(self.samp_sender[off:off + M]), ptr(logL), M, -1, ptr(si),
                                                  ptr(nl), ptr(ws_c), ws_c.numel(), s), "hipns_count_crossed_edges")
            torch.index_select(logL, 0, si[:M], out=lL[:M])
            nld[:M].copy_(nl[:M])
            check(lib().hipns_evidence_scan(ptr(lL), ptr(nld), M, -1, None, ptr(fins[r]), ptr(ws_e), ws_e.numel(),
                                            s), "hipns_evidence_scan")
            phs[r] = self.samp_phantom[off:off + M].sum()
        fin = fins.cpu().numpy()
        ph = phs.cpu().numpy()
        out = []
        for r in range(self.R):
            mu, var = _linear_to_log_stats(fin[r, 3], fin[r, 5])
            pf = ph[r] / Ms[r]
            k = pf / (1.0 - pf)
            out.append((mu, math.sqrt(var) * math.sqrt(1.0 + k)))
        return np.array(out), fin, ph

    def results_core(self, r: int, stream=None, full: bool = False, trim: bool = True) -> dict:
        """_to_results (sharded_static.py:652-745) for run r on the device: count_crossed_edges, the
        evidence scan and, with full=True, the per-sample finals (hipns_results_finals,
        hipns_prior_finals).  trim=True works on the first num_samples entries of the sample collection;
        trim=False on the whole collection (max_samples entries) with the dynamic forms that stop at
        num_samples (count_crossed_edges(num_samples=...), compute_evidence_stats(num_samples=...)),
        entries past it being the collection's placeholders."""
        torch = _torch()
        st = self.states()[r]
        spec = self.specs[r]
        d = self.runs_host[r]
        M = int(min(st.num_samples, d.max_samples))
        L = M if trim else int(d.max_samples)
        stop = -1 if trim else M
        off = int(d.samp_off)
        dev = self.device
        s = self._s(stream)
        logL = self.samp_logL[off:off + L]
        sender = self.samp_sender[off:off + L]
        si = torch.empty(L, dtype=torch.int64, device=dev)
        nl = torch.empty(L, dtype=torch.int64, device=dev)
        ws = torch.empty(int(lib().hipns_count_crossed_edges_workspace_size(L)), dtype=torch.uint8, device=dev)
        check(lib().hipns_count_crossed_edges(ptr(sender), ptr(logL), L, stop, ptr(si), ptr(nl), ptr(ws), ws.numel(),
                                              s), "hipns_count_crossed_edges")
        lL = logL.index_select(0, si)
        nld = nl.to(torch.float64)
        per8 = torch.empty((8, L), dtype=torch.float64, device=dev)
        fin8 = torch.empty(8, dtype=torch.float64, device=dev)
        ws2 = torch.empty(int(lib().hipns_evidence_scan_workspace_size(L)), dtype=torch.uint8, device=dev)
        check(lib().hipns_evidence_scan(ptr(lL), ptr(nld), L, stop, ptr(per8), ptr(fin8), ptr(ws2), ws2.numel(), s),
              "hipns_evidence_scan")
        fin = fin8.cpu().numpy()
        log_Z_mean, log_Z_var = _linear_to_log_stats(fin[3], fin[5])
        total_phantom = int(self.samp_phantom[off:off + M].to(torch.int64).sum().item())
        pf = total_phantom / M
        k = pf / (1.0 - pf)
        log_Z_uncert = math.sqrt(log_Z_var) * math.sqrt(1.0 + k)
        ess = math.exp((fin[3] + fin[3]) - fin[7]) / (1.0 + k)
        out = dict(log_Z_mean=float(log_Z_mean), log_Z_uncert=float(log_Z_uncert), ESS=float(ess), M=M, L=L,
                   total_phantom=total_phantom, termination_reason=int(st.termination_reason),
                   iterations=int(st.iteration), cap_hits=int(st.cap_hits))
        if self.refine_fp32:
            out["refined_below_contour"] = int(self.refine_below[r].item())
        if full:
            model = spec.model
            D = model.U_ndims
            ev = self.samp_evals[off:off + L].index_select(0, si)
            log_dp = torch.empty(L, dtype=torch.float64, device=dev)
            H = torch.empty(4, dtype=torch.float64, device=dev)
            ws3 = torch.empty(int(lib().hipns_results_finals_workspace_size(L)), dtype=torch.uint8, device=dev)
            check(lib().hipns_results_finals(ptr(per8[6]), ptr(lL), ptr(per8[1]), L, float(log_Z_mean), ptr(log_dp),
                                             ptr(H), ptr(ws3), ws3.numel(), s), "hipns_results_finals")
            U = torch.empty((L, D), dtype=torch.float64, device=dev)
            if model.dynamic_prior:
                # a prior program that depends on earlier prior values: the model's generator forms X
                check(lib().hipns_prior_finals(ptr(self.samp_U) + off * self.dps * 4, self.dps, ptr(si), M, L, D,
                                               *self._prior_program_dev(model), None, ptr(U), None, None, s),
                      "hipns_prior_finals")
                Uh = U.cpu().numpy()
                X = None
                log_post = lL.cpu().numpy() + model.log_prob_prior_numpy(Uh)
            else:
                X = torch.empty((L, D), dtype=torch.float64, device=dev)
                lpost = torch.empty(L, dtype=torch.float64, device=dev)
                check(lib().hipns_prior_finals(ptr(self.samp_U) + off * self.dps * 4, self.dps, ptr(si), M, L, D,
                                               *self._prior_program_dev(model), ptr(lL), ptr(U), ptr(X), ptr(lpost),
                                               s), "hipns_prior_finals")
                Uh = U.cpu().numpy()
                X = X.cpu().numpy()
                log_post = lpost.cpu().numpy()
            Hh = H.cpu().numpy()
            out.update(samples_indices=si.cpu().numpy(), num_live_points=nl.cpu().numpy(), log_L=lL.cpu().numpy(),
                       per8=per8.cpu().numpy(), evals=ev.to(torch.int64).cpu().numpy(), U=Uh, X=X,
                       log_dp=log_dp.cpu().numpy(), H_unstable=float(Hh[0]), H_stable=float(Hh[1]),
                       H_mean=float(Hh[2]), log_posterior_density=log_post)
        return out

    def _prior_program_dev(self, model):
        """(loc, S, qkind, diag) of the model's affine prior program as device arrays (cached per model)."""
        torch = _torch()
        cache = self.__dict__.setdefault("_prior_dev", {})
        key = id(model)
        if key not in cache:
            qkind, loc, S = model.prior_program()
            diag = int(np.allclose(S, np.diag(np.diag(S))))
            t = (torch.as_tensor(np.ascontiguousarray(loc), dtype=torch.float64, device=self.device),
                 torch.as_tensor(np.ascontiguousarray(S), dtype=torch.float64, device=self.device),
                 torch.as_tensor(np.ascontiguousarray(qkind), dtype=torch.int32, device=self.device), diag)
            cache[key] = (model, t)
        _, (loc, S, qk, diag) = cache[key]
        return ptr(loc), ptr(S), ptr(qk), diag

    def to_results(self, r: int, stream=None, trim: bool = True) -> NestedSamplerResults:
        """NestedSamplerResults of run r (sharded_static.py:652-773), every per-sample field formed on the
        device; trim=False keeps the whole collection (max_samples rows, placeholders past num_samples)."""
        core = self.results_core(r, stream=stream, full=True, trim=trim)
        if core["cap_hits"]:
            warnings.warn(f"run {r}: {core['cap_hits']} slices reached the step cap of {self.launch.max_steps} shrinks "
                          f"and kept their seed point; raise max_steps", RuntimeWarning)
        model = self.specs[r].model
        U = core["U"]
        samples = model.transform(U) if core["X"] is None else model.split_transformed(core["X"])
        evals = core["evals"]
        total_evals = int(np.sum(evals))
        M = core["M"]
        return NestedSamplerResults(
            log_Z_mean=core["log_Z_mean"], log_Z_uncert=core["log_Z_uncert"], ESS=core["ESS"], H_mean=core["H_mean"],
            samples=samples, parametrised_samples={}, U_samples=U, log_L_samples=core["log_L"],
            log_dp_mean=core["log_dp"], log_X_mean=core["per8"][1],
            log_posterior_density=core["log_posterior_density"], num_live_points_per_sample=core["num_live_points"],
            num_likelihood_evaluations_per_sample=evals, total_num_samples=M,
            total_phantom_samples=core["total_phantom"], total_num_likelihood_evaluations=total_evals,
            log_efficiency=float(np.log(M) - np.log(total_evals)), termination_reason=core["termination_reason"])


def run_queue(specs: List[RunSpec], slots: int, term: dict, *, graph_iters: int = 8,
              init_efficiency_threshold: float = 0.1, **batch_kw) -> List[dict]:
    """Runs of one likelihood family streamed through `slots` device slots: each slot holds one run;
    when a run terminates its live points are consumed, its log Z is formed (with the other runs that
    finished in the same interval) and the slot takes the next run of the queue.  The slots advance
    together through replays of a captured graph of `graph_iters` shells shaped for the slot capacity
    (the largest num_live_points, max_samples, dimension, num_slices and k of the queue); a slot whose
    run is done is skipped by every kernel, so runs of uneven length no longer hold the others back
    beyond one graph interval, and the GPU stays full until the queue drains.  Each run performs the
    same shells as in a batch of the same launch shape: its results do not depend on the slot or on
    the other runs.  Returns, per spec in order, log_Z_mean, log_Z_uncert, termination_reason,
    num_likelihood_evaluations, num_samples, iterations and cap_hits."""
    if not specs:
        return []
    if graph_iters <= 0:
        raise ValueError("run_queue needs graph_iters > 0")
    cap = dict(N=max(s.num_live_points for s in specs), max_samples=max(s.max_samples for s in specs),
               D=max(s.model.U_ndims for s in specs), num_slices=max(s.num_slices for s in specs),
               k_phantom=max(s.k_phantom for s in specs))
    S = min(int(slots), len(specs))
    b = DeviceBatch(list(specs[:S]), slot_capacity=cap, **batch_kw)
    owner = list(range(S))
    nxt = S
    results: List[Optional[dict]] = [None] * len(specs)
    # the uniform phase's check on a new run's register (efficiency 0) always stops it before a shell,
    # so every slot starts in the slice phase (init_slot)
    for r in range(S):
        b.init_slot(r, term, init_efficiency_threshold)
    g = b._get_graph(term, graph_iters)
    active = set(range(S))
    while active:
        flags = b.done_flags()
        fin = sorted(r for r in active if flags[r])
        if not fin:
            g.replay()
            continue
        for r in fin:
            b.consume_slot(r)
        z = b.log_Z_slots(fin)
        sts = b.states()
        for i, r in enumerate(fin):
            st = sts[r]
            results[owner[r]] = dict(log_Z_mean=float(z[i, 0]), log_Z_uncert=float(z[i, 1]),
                                     termination_reason=int(st.termination_reason),
                                     num_likelihood_evaluations=int(st.num_likelihood_evaluations),
                                     num_samples=int(st.num_samples), iterations=int(st.iteration),
                                     cap_hits=int(st.cap_hits))
        for r in fin:
            if nxt < len(specs):
                b.assign(r, specs[nxt])
                owner[r] = nxt
                nxt += 1
                b.init_slot(r, term, init_efficiency_threshold)
            else:
                active.discard(r)
    _torch().cuda.synchronize()
    return results


def run_concurrent(batches: List["DeviceBatch"], term, *, init_efficiency_threshold: float = 0.1,
                   graph_iters: int = 8, max_iterations: int = 1_000_000) -> None:
    """DeviceBatch.run for several batches at once, each on a stream of its own (default priority).
    The uniform-sampler phase goes shell by shell and then each batch's captured graph of `graph_iters`
    shells is replayed on its stream; in both the host cycles through the batches, reading one batch's
    done flags (a wait on that stream only) while the others' shells run, so one batch's launches fill
    the GPU while another's drain.  Every batch performs the same shells, in the same
    order, as batch.run(term, graph_iters=graph_iters) would: its results are identical.  `term` is
    one termination dict for all batches or a list of one per batch."""
    torch = _torch()
    terms = list(term) if isinstance(term, (list, tuple)) else [term] * len(batches)
    if any(b.external for b in batches):
        raise ValueError("run_concurrent: batches with a callable likelihood step on the host; use run()")
    if graph_iters <= 0:
        raise ValueError("run_concurrent needs graph_iters > 0")
    streams = [torch.cuda.Stream() for _ in batches]
    # the side streams do not wait for the default stream by themselves: DeviceBatch.__init__ fills
    # the live set, new chains and run states there, and every batch's first kernel must follow that
    cur = torch.cuda.current_stream()
    for s in streams:
        s.wait_stream(cur)
    uterms =[dict(efficiency_threshold=float(init_efficiency_threshold), dlogZ=0.0,
                   max_samples=float(min(sp.max_samples for sp in b.specs))) for b in batches]
    for b, s, ut in zip(batches, streams, uterms):
        with torch.cuda.stream(s):
            b.init_state()
            if init_efficiency_threshold > 0:
                b.check_termination(ut, True)
    # the uniform-sampler phase shell by shell, the batches in turn (one batch's shell runs while the
    # host waits for another's done flags)
    live = []
    if init_efficiency_threshold > 0:
        for i, (b, s) in enumerate(zip(batches, streams)):
            with torch.cuda.stream(s):
                if not np.all(b.done_flags()):
                    b.set_sampler(1)
                    live.append(i)
    its = [0] * len(batches)
    while live:
        nxt = []
        for i in live:
            b = batches[i]
            with torch.cuda.stream(streams[i]):
                if its[i] >= max_iterations or np.all(b.done_flags()):
                    b.set_sampler(b.slice_sampler)
                    continue
                b.collect_shell(uterms[i])
            its[i] += 1
            nxt.append(i)
        live = nxt
    for b, s, term in zip(batches, streams, terms):
        with torch.cuda.stream(s):
            b.check_termination(term, True)
    graphs = [b._get_graph(t, graph_iters) for b, t in zip(batches, terms)]  # each capture synchronises
    its = [0] * len(batches)
    live = list(range(len(batches)))
    while live:
        nxt = []
        for i in live:
            with torch.cuda.stream(streams[i]):
                if its[i] >= max_iterations or np.all(batches[i].done_flags()):
                    continue
                graphs[i].replay()
            its[i] += graph_iters
            nxt.append(i)
        live = nxt
    for b, s in zip(batches, streams):
        with torch.cuda.stream(s):
            b.consume()
    torch.cuda.synchronize()


def _linear_to_log_stats(log_f_mean, log_f2_mean):
    """internals/stats.py:55-74."""
    fsq = log_f_mean + log_f_mean
    mu = fsq - 0.5 * log_f2_mean
    s2 = log_f2_mean - fsq
    return mu, (s2 if np.isnan(s2) else max(s2, np.finfo(np.float64).eps))
