"""ctypes binding of libhipns.so (the C ABI declared in include/hipns.h).

The library is built in-tree by ``jaxns_amd/csrc/Makefile`` (``__graft_entry__.build()``).
There is no fallback: if the shared object is missing or a call fails, a
``HipnsError`` is raised.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhipns.so")
CSRC = os.path.join(_HERE, "csrc")

HIPNS_OK = 0
ERRORS = {-1: "EINVAL", -2: "EHIP", -3: "ENOMEM", -4: "ERANGE", -5: "ENOTSUP"}

FAM_MVN, FAM_GMM, FAM_LINREG = 0, 1, 2
MAX_DIM = 128
MAX_LIVE = 8192
MAX_GMM_K = 8


class HipnsError(RuntimeError):
    pass


P = C.POINTER
_i32, _i64, _u32, _d, _f = C.c_int32, C.c_int64, C.c_uint32, C.c_double, C.c_float
_vp = C.c_void_p


class RunDesc(C.Structure):
    _fields_ = [("D", _i32), ("N", _i32), ("shell", _i32), ("num_slices", _i32), ("k_phantom", _i32),
                ("midpoint_shrink", _i32), ("sampler", _i32), ("max_uniform_evals", _i32),
                ("max_samples", _i64), ("live_off", _i64), ("samp_off", _i64), ("chain_off", _i64),
                ("param_off", _i64), ("seed_lo", _u32), ("seed_hi", _u32), ("log_const", _d)]


class RunState(C.Structure):
    _fields_ = [("next_idx", _i64), ("num_samples", _i64), ("iteration", _i64),
                ("num_likelihood_evaluations", _i64), ("termination_reason", _i64), ("num_samples_used", _i64),
                ("evidence_calc", _d * 8), ("evidence_calc_with_remaining", _d * 8), ("log_L_contour", _d),
                ("efficiency", _d), ("relative_spread", _d), ("absolute_spread", _d), ("peak_log_XL", _d),
                ("contour", _d), ("plateau", _i32), ("no_seed_points", _i32), ("done", _i32), ("buf", _i32),
                ("first_seed", _i32), ("sender_node", _i32), ("flags", _i32), ("pad0", _i32), ("cap_hits", _i64)]


_TC_NAMES = ["ess", "evidence_uncert", "dlogZ", "max_samples", "max_num_likelihood_evaluations", "log_L_contour",
             "efficiency_threshold", "rtol", "atol", "peak_XL_frac"]


class TermCond(C.Structure):
    _fields_ = [("has_" + n, _i32) for n in _TC_NAMES] + [(n, _d) for n in _TC_NAMES]


class Buffers(C.Structure):
    _fields_ = [("live_U", _vp * 2), ("live_logL", _vp * 2), ("live_cons", _vp * 2), ("live_sender", _vp * 2),
                ("live_evals", _vp * 2), ("samp_U", _vp), ("samp_logL", _vp), ("samp_sender", _vp),
                ("samp_evals", _vp), ("samp_phantom", _vp), ("new_U", _vp), ("new_logL", _vp), ("new_evals", _vp),
                ("params", _vp), ("dps", _i32), ("pad0", _i32), ("run_order", _vp)]


class Launch(C.Structure):
    _fields_ = [("family", _i32), ("lanes_per_chain", _i32), ("waves_per_block", _i32), ("gmm_K", _i32),
                ("lr_ndata", _i32), ("max_steps", _i32), ("lr_panels", _i32), ("lr_terms", _i32),
                ("lr_data_off", _i64)]


class ExtState(C.Structure):
    _fields_ = [("U0", _vp), ("nv", _vp), ("prop_U", _vp), ("logL0", _vp), ("prop_logL", _vp), ("left", _vp),
                ("right", _vp), ("t", _vp), ("s", _vp), ("k", _vp), ("evals", _vp), ("flags", _vp), ("grad", _vp)]


assert C.sizeof(RunDesc) == 88, C.sizeof(RunDesc)
assert C.sizeof(Launch) == 40, C.sizeof(Launch)

_lib = None


def build(force: bool = False, jobs: int = 8) -> str:
    """Compile libhipns.so for gfx950 with hipcc (cross-compiles without a GPU)."""
    args = ["make", "-C", CSRC, f"-j{jobs}"]
    if force:
        subprocess.run(["make", "-C", CSRC, "clean"], check=True, capture_output=True)
    subprocess.run(args, check=True)
    return LIB_PATH


ABI_VERSION = 1  # HIPNS_ABI_VERSION of include/hipns.h that these struct mirrors follow


def lib():
    """Load libhipns.so; raises HipnsError when it has not been built or is of another ABI version."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HipnsError(f"{LIB_PATH} is missing: build it with `make -C {CSRC}` (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        _declare(L)
        if L.hipns_abi_version() != ABI_VERSION:
            raise HipnsError(f"{LIB_PATH} has ABI version {L.hipns_abi_version()}, these bindings need "
                             f"{ABI_VERSION}: rebuild it with `make -C {CSRC}`")
        _lib = L
    return _lib


def _declare(L):
    L.hipns_abi_version.restype = C.c_int
    L.hipns_last_error.argtypes = [C.c_char_p, C.c_size_t]
    L.hipns_create_init_state.argtypes = [_vp, _vp, _i32, P(RunDesc), P(Buffers), P(Launch), _vp, C.c_size_t, _vp]
    L.hipns_collect_shell.argtypes = [_vp, _vp, _i32, P(RunDesc), P(Buffers), P(Launch), P(TermCond), _vp,
                                      C.c_size_t, _vp]
    L.hipns_consume_live.argtypes = [_vp, _vp, _i32, P(RunDesc), P(Buffers), _vp]
    L.hipns_shell_discard.argtypes = [_vp, _vp, _i32, P(RunDesc), P(Buffers), _vp]
    L.hipns_shell_sample.argtypes = [_vp, _vp, _i32, P(RunDesc), P(Buffers), P(Launch), _vp]
    L.hipns_shell_replace.argtypes = [_vp, _vp, _i32, P(RunDesc), P(Buffers), _vp]
    L.hipns_shell_register.argtypes = [_vp, _vp, _i32, P(RunDesc), P(Buffers), P(TermCond), _vp]
    L.hipns_check_termination.argtypes = [_vp, _vp, _i32, P(Buffers), P(TermCond), _i32, _vp]
    L.hipns_collect_shell_workspace_size.restype = C.c_size_t
    L.hipns_collect_shell_workspace_size.argtypes = [_i32, P(RunDesc)]
    L.hipns_evidence_scan.argtypes = [_vp, _vp, _i64, _i64, _vp, _vp, _vp, C.c_size_t, _vp]
    L.hipns_evidence_scan_workspace_size.restype = C.c_size_t
    L.hipns_evidence_scan_workspace_size.argtypes = [_i64]
    L.hipns_count_crossed_edges.argtypes = [_vp, _vp, _i64, _i64, _vp, _vp, _vp, C.c_size_t, _vp]
    L.hipns_count_crossed_edges_workspace_size.restype = C.c_size_t
    L.hipns_count_crossed_edges_workspace_size.argtypes = [_i64]
    L.hipns_argsort_stable.argtypes = [_vp, _i64, _vp, _vp, C.c_size_t, _vp]
    L.hipns_argsort_workspace_size.restype = C.c_size_t
    L.hipns_argsort_workspace_size.argtypes = [_i64]
    L.hipns_log_likelihood.argtypes = [_vp, _i64, _vp, _vp, P(RunDesc), P(Buffers), P(Launch), _vp]
    L.hipns_log_likelihood_grad.argtypes = [_vp, _i64, _vp, _vp, _vp, P(RunDesc), P(Buffers), P(Launch), _vp]
    L.hipns_ext_step.argtypes = [_vp, _vp, _i32, P(RunDesc), P(Buffers), P(Launch), P(ExtState), _i32, _i32, _vp]
    L.hipns_init_prepare.argtypes = [_vp, _vp, _i32, P(RunDesc), P(Buffers), _vp]
    L.hipns_init_finalize.argtypes = [_vp, _vp, _i32, P(RunDesc), P(Buffers), _vp]
    L.hipns_batch_evidence.argtypes = [_vp, _vp, _i32, P(RunDesc), P(Buffers), _vp, _vp, _vp, C.c_size_t, _vp]
    L.hipns_batch_evidence_workspace_size.restype = C.c_size_t
    L.hipns_batch_evidence_workspace_size.argtypes = [_i32, P(RunDesc)]
    L.hipns_linreg_stats.argtypes = [_vp, _i64, _vp, _vp, _vp, _vp, P(RunDesc), P(Buffers), P(Launch), _vp]
    L.hipns_prior_transform.argtypes = [_vp, _i64, _i32, _i32, _vp, _vp, _vp, _vp, _vp, _i32, _vp]
    L.hipns_slice_chains_injected.argtypes = [_vp, P(RunDesc), P(Buffers), P(Launch), _i32, _i32, _d, _vp, _vp,
                                              _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    L.hipns_philox4x32.argtypes = [_vp, _i64, _u32, _u32, _vp, _vp]
    L.hipns_ndtri_f32.argtypes = [_vp, _i64, _vp, _vp]
    L.hipns_slice_quotients.argtypes = [_vp, _vp, _i64, _vp, _vp]
    L.hipns_cholesky_f64.argtypes = [_vp, _i32, _vp, _vp, _vp]
    L.hipns_mvn_prepare_f64.argtypes = [_vp, _vp, _vp, _vp, _vp, _i32, _i32, _vp, _vp, _vp, _vp, _vp, C.c_size_t,
                                        _vp]
    L.hipns_mvn_prepare_workspace_size.restype = C.c_size_t
    L.hipns_mvn_prepare_workspace_size.argtypes = [_i32]
    L.hipns_evidence_scan_noalloc.argtypes = [_vp, _vp, _i64, _i64, _vp, _vp, _vp]
    L.hipns_evidence_scan_noalloc_max.restype = C.c_size_t
    L.hipns_evidence_scan_noalloc_max.argtypes = []
    L.hipns_results_finals.argtypes = [_vp, _vp, _vp, _i64, _d, _vp, _vp, _vp, C.c_size_t, _vp]
    L.hipns_results_finals_workspace_size.restype = C.c_size_t
    L.hipns_results_finals_workspace_size.argtypes = [_i64]
    L.hipns_prior_finals.argtypes = [_vp, _i32, _vp, _i64, _i64, _i32, _vp, _vp, _vp, _i32, _vp, _vp, _vp, _vp, _vp]
    L.hipns_seed_indices.argtypes = [_u32, _u32, _u32, _i32, _i32, _i32, _vp, _vp, _vp]
    L.hipns_slice_directions.argtypes = [_u32, _u32, _u32, _i32, _i32, _i32, _i32, _vp, _vp]


EXPORTED = ["hipns_abi_version", "hipns_last_error", "hipns_create_init_state", "hipns_collect_shell",
            "hipns_shell_discard", "hipns_shell_sample", "hipns_shell_replace", "hipns_shell_register",
            "hipns_consume_live", "hipns_check_termination", "hipns_collect_shell_workspace_size",
            "hipns_evidence_scan", "hipns_evidence_scan_workspace_size", "hipns_count_crossed_edges",
            "hipns_count_crossed_edges_workspace_size", "hipns_argsort_stable", "hipns_argsort_workspace_size",
            "hipns_log_likelihood", "hipns_log_likelihood_grad", "hipns_linreg_stats", "hipns_batch_evidence",
            "hipns_batch_evidence_workspace_size", "hipns_ext_step", "hipns_init_prepare", "hipns_init_finalize",
            "hipns_prior_transform", "hipns_slice_chains_injected", "hipns_philox4x32",
            "hipns_ndtri_f32", "hipns_slice_quotients", "hipns_seed_indices", "hipns_slice_directions",
            "hipns_results_finals", "hipns_results_finals_workspace_size", "hipns_prior_finals",
            "hipns_evidence_scan_noalloc", "hipns_evidence_scan_noalloc_max", "hipns_cholesky_f64",
            "hipns_mvn_prepare_f64", "hipns_mvn_prepare_workspace_size"]


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    lib().hipns_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(rc: int, what: str = "hipns call") -> None:
    if rc != HIPNS_OK:
        raise HipnsError(f"{what} failed with {ERRORS.get(rc, rc)}: {last_error()}")


def ptr(t) -> int:
    """Device pointer of a torch tensor (or None -> NULL)."""
    if t is None:
        return None
    return t.data_ptr()


def stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def make_term_cond(**kw) -> TermCond:
    tc = TermCond()
    for k, v in kw.items():
        if v is None:
            continue
        setattr(tc, "has_" + k, 1)
        setattr(tc, k, float(v))
    return tc
