"""CPU oracle for the jaxns nested-sampling hot path -- TEST INFRASTRUCTURE.

This package is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
it.  It wraps ``oracle/_build/libhipns_oracle.so`` (built from
``oracle/hipns_oracle.c`` by ``oracle/Makefile``), a plain-C restatement of
the reference algorithm; each C function cites the jaxns file:line it follows.

Parity status of each oracle function (how it is pinned, see tests/test_oracle.py):
  logaddexp / signed_logaddexp / cumulative_logsumexp -- reference KATs
      (test_log_semiring.py:8-67) and numpy.logaddexp.
  evidence_update / compute_evidence_stats -- closed forms (X_i = prod n/(n+1),
      constant-L evidence, E[Z] of a uniform-shrinkage model), brute-force
      linear-space recurrence, dynamic-stop semantics of test_cumulative_ops.py.
  linear_to_log_stats -- test_stats.py:7-16 log-normal check.
  argsort_stable -- numpy.argsort(kind='stable').
  count_crossed_edges -- test_tree_structure.py:19-80 KATs and the O(N^2)
      brute force count_intervals_naive on random trees.
  seed_index -- brute-force linear search over the masked set.
  update_start / replace_index -- the reference's test_maps.py:7-30 known answers.
  ndtri -- scipy.special.ndtri.
  philox4x32 -- Random123 published known-answer vectors.
  slice chain -- invariants (bracket on the cube boundary, t inside the bracket,
      acceptance rule, alpha=1 reduction), a 1-D constant-likelihood KS test and a
      D-ball radius^D KS test.
  ns_run -- analytic log Z of the benchmark models.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libhipns_oracle.so")

_lib = None


def build() -> str:
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


P = C.POINTER
_d = C.c_double
_i64 = C.c_int64
_i32 = C.c_int32


class Model(C.Structure):
    _fields_ = [
        ("family", C.c_int), ("D", C.c_int), ("prior_kind", C.c_int),
        ("prior_loc", P(_d)), ("prior_scale", P(_d)), ("prior_tril", P(_d)),
        ("mvn_mu", P(_d)), ("mvn_chol", P(_d)),
        ("K", C.c_int), ("gmm_means", P(_d)), ("gmm_sigma", P(_d)), ("gmm_logw", P(_d)),
        ("Ndata", _i64), ("lr_X", P(_d)), ("lr_y", P(_d)), ("lr_sigma", _d),
    ]


class TermCond(C.Structure):
    _fields_ = []
    for _n in ["ess", "evidence_uncert", "dlogZ", "max_samples", "max_num_likelihood_evaluations",
               "log_L_contour", "efficiency_threshold", "rtol", "atol", "peak_XL_frac"]:
        _fields_ += [("has_" + _n, C.c_int), (_n, _d)]


class Register(C.Structure):
    _fields_ = [("num_samples_used", _i64), ("evidence_calc", _d * 8), ("evidence_calc_with_remaining", _d * 8),
                ("num_likelihood_evaluations", _i64), ("log_L_contour", _d), ("efficiency", _d),
                ("plateau", _i32), ("no_seed_points", _i32), ("relative_spread", _d),
                ("absolute_spread", _d), ("peak_log_XL", _d)]


class RunCfg(C.Structure):
    _fields_ = [("num_live_points", C.c_int), ("num_slices", C.c_int), ("k_phantom", C.c_int),
                ("midpoint_shrink", C.c_int), ("max_samples", _i64), ("init_efficiency_threshold", _d),
                ("seed", C.c_uint64), ("num_threads", C.c_int), ("max_iterations", _i64), ("dir16", C.c_int)]


class RunResult(C.Structure):
    _fields_ = [("termination_reason", _i64), ("num_samples", _i64), ("num_iterations", _i64),
                ("total_likelihood_evaluations", _i64), ("log_Z_mean", _d), ("log_Z_uncert", _d),
                ("ESS", _d), ("H_mean", _d), ("log_efficiency", _d), ("total_phantom_samples", _i64)]


FAM_MVN, FAM_GMM, FAM_LINREG = 0, 1, 2
PRIOR_NORMAL, PRIOR_UNIFORM, PRIOR_MVN = 0, 1, 2


def _declare(L):
    L.orc_logaddexp.restype = _d
    L.orc_logaddexp.argtypes = [_d, _d]
    L.orc_signed_logaddexp.argtypes = [_d, _d, _d, _d, P(_d), P(_d)]
    L.orc_cumulative_logsumexp.argtypes = [P(_d), _i64, P(_d)]
    L.orc_evidence_init.argtypes = [P(_d)]
    L.orc_evidence_update.argtypes = [P(_d), _d, _d]
    L.orc_compute_evidence_stats.argtypes = [P(_d), P(_d), _i64, _i64, P(_d), P(_d)]
    L.orc_linear_to_log_stats.argtypes = [_d, _d, P(_d), P(_d)]
    L.orc_ess_kish.restype = _d
    L.orc_ess_kish.argtypes = [_d, _d]
    L.orc_argsort_stable.argtypes = [P(_d), _i64, P(_i64)]
    L.orc_count_crossed_edges.argtypes = [P(_i64), P(_d), _i64, _i64, P(_i64), P(_i64)]
    L.orc_count_intervals_naive.argtypes = [P(_i64), P(_d), _i64, P(_i64), P(_i64)]
    L.orc_update_start.restype = _i64
    L.orc_update_start.argtypes = [_i64, _i64, _i64]
    L.orc_seed_index.restype = _i64
    L.orc_seed_index.argtypes = [P(_d), _i64, _d, _d]
    L.orc_ndtri.restype = _d
    L.orc_ndtri.argtypes = [_d]
    L.orc_ndtr.restype = _d
    L.orc_ndtr.argtypes = [_d]
    L.orc_philox4x32.argtypes = [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]
    L.orc_prior_transform.argtypes = [P(Model), P(_d), P(_d)]
    L.orc_log_likelihood.restype = _d
    L.orc_log_likelihood.argtypes = [P(Model), P(_d)]
    L.orc_log_prob_prior.restype = _d
    L.orc_log_prob_prior.argtypes = [P(Model), P(_d)]
    for sfx, R in (("f64", _d), ("f32", C.c_float)):
        f = getattr(L, "orc_slice_chain_" + sfx)
        f.restype = C.c_int
        f.argtypes = [P(Model), P(R), _d, _d, C.c_int, C.c_int, C.c_int, P(R), P(R), C.c_int, P(_d),
                      P(R), P(_d), P(_i64), P(R), P(_d), P(R), P(_d), P(_i32)]
        g = getattr(L, "orc_slice_bounds_" + sfx)
        g.argtypes = [P(R), P(R), C.c_int, P(R), P(R)]
    L.orc_init_register.argtypes = [P(Register)]
    L.orc_determine_termination.restype = C.c_int
    L.orc_determine_termination.argtypes = [P(TermCond), P(Register), P(_i64)]
    L.orc_ns_run.restype = C.c_int
    L.orc_ns_run.argtypes = [P(Model), P(RunCfg), P(TermCond), P(_d), P(_d), P(_i64), P(_i64), P(C.c_int8),
                             P(RunResult)]
    L.orc_to_results.restype = C.c_int
    L.orc_to_results.argtypes = [P(_d), P(_i64), P(_i64), P(C.c_int8), _i64, P(_i64), P(_i64), P(_d),
                                 P(RunResult)]


def _p(a, ct):
    return a.ctypes.data_as(P(ct))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i64a(a):
    return np.ascontiguousarray(a, dtype=np.int64)


# ---------------------------------------------------------------- wrappers
def logaddexp(a: float, b: float) -> float:
    return lib().orc_logaddexp(a, b)


def signed_logaddexp(a, sa, b, sb):
    o, s = _d(), _d()
    lib().orc_signed_logaddexp(a, sa, b, sb, C.byref(o), C.byref(s))
    return o.value, s.value


def cumulative_logsumexp(u):
    u = _f64(u)
    out = np.empty_like(u)
    lib().orc_cumulative_logsumexp(_p(u, _d), u.size, _p(out, _d))
    return out


EVIDENCE_FIELDS = ("log_L", "log_X_mean", "log_X2_mean", "log_Z_mean", "log_ZX_mean", "log_Z2_mean",
                   "log_dZ_mean", "log_dZ2_mean")


def evidence_init():
    s = np.empty(8)
    lib().orc_evidence_init(_p(s, _d))
    return s


def evidence_update(state, n, log_L_next):
    s = _f64(state).copy()
    lib().orc_evidence_update(_p(s, _d), float(n), float(log_L_next))
    return s


def compute_evidence_stats(log_L, num_live_points, stop_idx=None):
    log_L = _f64(log_L)
    n = _f64(num_live_points)
    M = log_L.size
    fin = np.empty(8)
    per = np.empty((M, 8))
    lib().orc_compute_evidence_stats(_p(log_L, _d), _p(n, _d), M, -1 if stop_idx is None else int(stop_idx),
                                     _p(fin, _d), _p(per, _d))
    return fin, per


def linear_to_log_stats(log_f_mean, log_f2_mean):
    mu, var = _d(), _d()
    lib().orc_linear_to_log_stats(float(log_f_mean), float(log_f2_mean), C.byref(mu), C.byref(var))
    return mu.value, var.value


def ess_kish(log_Z_mean, log_dZ2_mean):
    return lib().orc_ess_kish(float(log_Z_mean), float(log_dZ2_mean))


def argsort_stable(keys):
    k = _f64(keys)
    out = np.empty(k.size, np.int64)
    lib().orc_argsort_stable(_p(k, _d), k.size, _p(out, _i64))
    return out


def count_crossed_edges(sender, log_L, num_samples=None):
    s = _i64a(sender)
    L = _f64(log_L)
    si = np.empty(s.size, np.int64)
    nl = np.empty(s.size, np.int64)
    lib().orc_count_crossed_edges(_p(s, _i64), _p(L, _d), s.size, -1 if num_samples is None else int(num_samples),
                                  _p(si, _i64), _p(nl, _i64))
    return si, nl


def count_intervals_naive(sender, log_L):
    s = _i64a(sender)
    L = _f64(log_L)
    si = np.empty(s.size, np.int64)
    nl = np.empty(s.size, np.int64)
    lib().orc_count_intervals_naive(_p(s, _i64), _p(L, _d), s.size, _p(si, _i64), _p(nl, _i64))
    return si, nl


def update_start(start, count, size):
    return int(lib().orc_update_start(int(start), int(count), int(size)))


def replace_index(operand, update, start):
    """replace_index (internals/maps.py:15-25) on host arrays, through the clamped start."""
    out = np.array(operand, copy=True)
    j = update_start(start, len(update), len(out))
    out[j:j + len(update)] = update
    return out


def seed_index(live_log_L, contour, u):
    L = _f64(live_log_L)
    return int(lib().orc_seed_index(_p(L, _d), L.size, float(contour), float(u)))


def ndtri(p):
    f = lib().orc_ndtri
    return np.array([f(float(x)) for x in np.ravel(p)]).reshape(np.shape(p))


def ndtr(x):
    f = lib().orc_ndtr
    return np.array([f(float(v)) for v in np.ravel(x)]).reshape(np.shape(x))


def philox4x32(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    o = np.empty(4, np.uint32)
    lib().orc_philox4x32(_p(c, C.c_uint32), _p(k, C.c_uint32), _p(o, C.c_uint32))
    return o


class ModelSpec:
    """Host-side model description handed to the C oracle (keeps arrays alive)."""

    def __init__(self, family, D, prior_kind, prior_loc, prior_scale, prior_tril=None, mvn_mu=None,
                 mvn_chol=None, gmm_means=None, gmm_sigma=None, gmm_logw=None, lr_X=None, lr_y=None,
                 lr_sigma=1.0):
        self._keep = []

        def arr(a):
            if a is None:
                return None
            a = _f64(a)
            self._keep.append(a)
            return _p(a, _d)

        self.D = int(D)
        self.s = Model()
        self.s.family = family
        self.s.D = D
        self.s.prior_kind = prior_kind
        self.s.prior_loc = arr(prior_loc)
        self.s.prior_scale = arr(prior_scale)
        self.s.prior_tril = arr(prior_tril)
        self.s.mvn_mu = arr(mvn_mu)
        self.s.mvn_chol = arr(mvn_chol)
        self.s.K = 0 if gmm_means is None else int(np.shape(gmm_means)[0])
        self.s.gmm_means = arr(gmm_means)
        self.s.gmm_sigma = arr(gmm_sigma)
        self.s.gmm_logw = arr(gmm_logw)
        self.s.Ndata = 0 if lr_X is None else int(np.shape(lr_X)[0])
        self.s.lr_X = arr(lr_X)
        self.s.lr_y = arr(lr_y)
        self.s.lr_sigma = float(lr_sigma)

    def log_likelihood(self, U):
        U = np.atleast_2d(_f64(U))
        f = lib().orc_log_likelihood
        return np.array([f(C.byref(self.s), _p(np.ascontiguousarray(u), _d)) for u in U])

    def prior_transform(self, U):
        U = np.atleast_2d(_f64(U))
        out = np.empty_like(U)
        for i, u in enumerate(U):
            u = np.ascontiguousarray(u)
            o = np.empty(self.D)
            lib().orc_prior_transform(C.byref(self.s), _p(u, _d), _p(o, _d))
            out[i] = o
        return out

    def log_prob_prior(self, U):
        U = np.atleast_2d(_f64(U))
        f = lib().orc_log_prob_prior
        return np.array([f(C.byref(self.s), _p(np.ascontiguousarray(u), _d)) for u in U])


def slice_chain(spec: ModelSpec, U0, logL0, contour, num_slices, dirs, tu, *, midpoint_shrink=True, k_phantom=0,
                dtype=np.float64, logL_replay=None):
    """One UniDimSliceSampler chain with injected draws; returns a dict of outputs and the trace."""
    D = spec.D
    R = _d if dtype == np.float64 else C.c_float
    fn = lib().orc_slice_chain_f64 if dtype == np.float64 else lib().orc_slice_chain_f32
    U0 = np.ascontiguousarray(U0, dtype=dtype)
    dirs = np.ascontiguousarray(dirs, dtype=dtype).reshape(num_slices, D)
    tu = np.ascontiguousarray(tu, dtype=dtype)
    max_steps = tu.shape[-1]
    tu = tu.reshape(num_slices, max_steps)
    rep = None if logL_replay is None else _f64(logL_replay).reshape(num_slices, max_steps)
    U_out = np.empty(D, dtype)
    logL_out = _d()
    ev = _i64()
    kk = max(k_phantom, 1)
    ph_U = np.zeros((kk, D), dtype)
    ph_L = np.zeros(kk)
    tr_t = np.full((num_slices, max_steps), np.nan, dtype)
    tr_L = np.full((num_slices, max_steps), np.nan)
    steps = np.zeros(num_slices, np.int32)
    rc = fn(C.byref(spec.s), _p(U0, R), float(logL0), float(contour), int(num_slices), int(midpoint_shrink),
            int(k_phantom), _p(dirs, R), _p(tu, R), int(max_steps), None if rep is None else _p(rep, _d),
            _p(U_out, R), C.byref(logL_out), C.byref(ev), _p(ph_U, R), _p(ph_L, _d), _p(tr_t, R), _p(tr_L, _d),
            _p(steps, _i32))
    if rc != 0:
        raise RuntimeError("oracle slice chain exceeded max_steps")
    return dict(U=U_out, logL=logL_out.value, evals=ev.value, phantom_U=ph_U[:k_phantom],
                phantom_logL=ph_L[:k_phantom], trace_t=tr_t, trace_logL=tr_L, steps=steps)


def slice_bounds(U0, n, dtype=np.float64):
    R = _d if dtype == np.float64 else C.c_float
    fn = lib().orc_slice_bounds_f64 if dtype == np.float64 else lib().orc_slice_bounds_f32
    U0 = np.ascontiguousarray(U0, dtype=dtype)
    n = np.ascontiguousarray(n, dtype=dtype)
    lft, rgt = R(), R()
    fn(_p(U0, R), _p(n, R), U0.size, C.byref(lft), C.byref(rgt))
    return lft.value, rgt.value


def make_term_cond(**kw) -> TermCond:
    tc = TermCond()
    for k, v in kw.items():
        if v is None:
            continue
        setattr(tc, "has_" + k, 1)
        setattr(tc, k, float(v))
    return tc


def determine_termination(term_kw: dict, reg: Register):
    tc = make_term_cond(**term_kw)
    why = _i64()
    done = lib().orc_determine_termination(C.byref(tc), C.byref(reg), C.byref(why))
    return bool(done), int(why.value)


def init_register() -> Register:
    r = Register()
    lib().orc_init_register(C.byref(r))
    return r


def ns_run(spec: ModelSpec, *, num_live_points, num_slices, k_phantom=0, midpoint_shrink=True, max_samples,
           seed=0, num_threads=0, max_iterations=-1, init_efficiency_threshold=0.1, term=None, dir16=False):
    """Full nested-sampling run (create_init_state -> phases -> to_results) on host cores.  dir16: slice
    directions from 16-bit uniform pairs as the device draws them, instead of FP64 Box-Muller."""
    cfg = RunCfg(num_live_points, num_slices, k_phantom, int(midpoint_shrink), int(max_samples),
                 float(init_efficiency_threshold), int(seed), int(num_threads), int(max_iterations), int(dir16))
    if term is None:
        term = dict(dlogZ=np.log(1.0 + 1e-3), max_samples=float(max_samples))
    tc = make_term_cond(**term)
    D = spec.D
    sU = np.empty((max_samples, D))
    sL = np.empty(max_samples)
    sS = np.empty(max_samples, np.int64)
    sE = np.empty(max_samples, np.int64)
    sP = np.empty(max_samples, np.int8)
    res = RunResult()
    rc = lib().orc_ns_run(C.byref(spec.s), C.byref(cfg), C.byref(tc), _p(sU, _d), _p(sL, _d), _p(sS, _i64),
                          _p(sE, _i64), _p(sP, C.c_int8), C.byref(res))
    if rc != 0:
        raise RuntimeError("oracle ns_run failed")
    out = {f: getattr(res, f) for f, _ in RunResult._fields_}
    M = min(res.num_samples, max_samples)
    out.update(samples_U=sU[:M], samples_logL=sL[:M], samples_sender=sS[:M], samples_evals=sE[:M],
               samples_phantom=sP[:M])
    return out


def to_results(samp_logL, samp_sender, samp_evals, samp_phantom):
    L = _f64(samp_logL)
    S = _i64a(samp_sender)
    E = _i64a(samp_evals)
    Ph = np.ascontiguousarray(samp_phantom, dtype=np.int8)
    M = L.size
    si = np.empty(M, np.int64)
    nl = np.empty(M, np.int64)
    per = np.empty((M, 8))
    res = RunResult()
    lib().orc_to_results(_p(L, _d), _p(S, _i64), _p(E, _i64), _p(Ph, C.c_int8), M, _p(si, _i64), _p(nl, _i64),
                         _p(per, _d), C.byref(res))
    out = {f: getattr(res, f) for f, _ in RunResult._fields_}
    out.update(samples_indices=si, num_live_points=nl, per_sample=per)
    return out
