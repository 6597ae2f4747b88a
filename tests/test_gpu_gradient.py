"""Gradient-guided slice sampling in the fused kernel (uni_slice_sampler.py:255-271).

After each accepted slice a gradient-guided chain spends one evaluation on d logL / dU at the accepted
point and reflects its last direction about it.  For the MVN and mixture families the sampling kernel
computes that gradient itself (sampler code 2): these tests check the device gradient against central
differences of the oracle's FP64 likelihood, and whole gradient-guided runs against analytic
evidences, next to the stepwise form whose gradient comes from torch autograd.
"""
import ctypes as C
import math

import numpy as np
import pytest

import jaxns_amd as J
from jaxns_amd import _lib
from jaxns_amd._lib import check, lib, ptr

from problems import gmm_oracle_spec, gmm_problem, mvn_oracle_spec, mvn_problem

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _fd_grad(spec, U, h=1e-6):
    """Central differences of the oracle's FP64 log-likelihood in each coordinate of U."""
    B, D = U.shape
    g = np.empty((B, D))
    for d in range(D):
        e = np.zeros(D)
        e[d] = h
        g[:, d] = (spec.log_likelihood(U + e) - spec.log_likelihood(U - e)) / (2 * h)
    return g


@pytest.mark.parametrize("fam,D,lpc", [("mvn", 6, 2), ("mvn", 6, 4), ("mvn", 40, 4), ("mvn", 20, 2),
                                       ("gmm", 5, 2), ("gmm", 12, 4), ("gmm", 40, 8)])
def test_loglik_grad_vs_finite_differences(fam, D, lpc):
    """hipns_log_likelihood_grad (the gradient the fused gradient-guided chains use) against central
    differences of the oracle's FP64 likelihood: every component within 2e-5 of the largest component
    of its row (FP32 products and the FP32 ndtri; about 1e-6 measured), and log L within 1e-5
    relative.  The MVN cases cover
    one and several 16/32-row output blocks of the second GEMM (W^T r); the mixture at D = 40 (row stride
    64) runs at 8 lanes per chain."""
    import torch
    from jaxns_amd.engine import DeviceBatch, RunSpec
    say(f"loglik grad {fam} D={D} lpc={lpc}")
    rng = np.random.default_rng(D + 10 * lpc)
    B = 300
    if fam == "mvn":
        model, _, mu, Sig = mvn_problem(D, seed=D)
        spec = mvn_oracle_spec(D, mu, Sig)
        U = rng.uniform(0.05, 0.95, size=(B, D))
    else:
        model, _, means, sig = gmm_problem(D, seed=D)
        spec = gmm_oracle_spec(D, means, sig)
        # points within a few widths of the components, where the gradient is not underflowed
        U = ((means[rng.integers(0, means.shape[0], B)] + rng.normal(size=(B, D)) * 0.4 + 10) / 20).clip(0.01, 0.99)
    U = U.astype(np.float32).astype(np.float64)
    b = DeviceBatch([RunSpec(model=model, num_live_points=64, num_slices=2 * D, max_samples=64 * 20, seed=1)],
                    lanes_per_chain=lpc, waves_per_block=2)
    Up = np.full((B, b.dps), 0.5, np.float32)
    Up[:, :D] = U
    Ud = torch.from_numpy(Up).cuda()
    out = torch.empty(B, dtype=torch.float64, device="cuda")
    gd = torch.zeros((B, b.dps), dtype=torch.float32, device="cuda")
    check(lib().hipns_log_likelihood_grad(ptr(Ud), B, ptr(out), ptr(gd), ptr(b.runs_dev), b.runs_host,
                                          C.byref(b.bufs), C.byref(b.launch), _lib.stream_ptr()))
    got_L = out.cpu().numpy()
    got = gd.cpu().numpy()[:, :D].astype(np.float64)
    ref_L = spec.log_likelihood(U)
    ref = _fd_grad(spec, U)
    scale = np.abs(ref).max(axis=1, keepdims=True)
    err = np.abs(got - ref) / np.maximum(scale, 1e-30)
    say(f"  max |dg| / max|g| per row {err.max():.2e}, log L rel {np.abs(got_L / ref_L - 1).max():.2e}")
    assert np.abs(got_L - ref_L).max() <= 1e-5 * np.abs(ref_L).max()
    assert err.max() < 2e-5
    # padding entries past D are not written
    assert np.all(gd.cpu().numpy()[:, D:] == 0)


@pytest.mark.parametrize("which,lpc", [("mvn", 2), ("gmm", 2), ("gmm8", 8)])
def test_gradient_guided_fused_logZ(which, lpc):
    """NestedSampler(gradient_guided=True) on a fused family runs the fused kernel (sampler code 2): log Z
    against the analytic value, no step-cap hits, and every slice-sampled point spent at least two
    evaluations per slice (its accepted proposal plus the gradient at it)."""
    say(f"gradient-guided fused {which} lpc={lpc}")
    if which == "mvn":
        model, log_Z, _, _ = mvn_problem(8, seed=4)
        ns = J.NestedSampler(model, num_live_points=200, s=3, max_samples=200 * 150, gradient_guided=True,
                             lanes_per_chain=lpc)
    elif which == "gmm":
        model, log_Z, _, _ = gmm_problem(3, seed=7)
        ns = J.NestedSampler(model, num_live_points=300, s=4, max_samples=300 * 200, gradient_guided=True,
                             lanes_per_chain=lpc)
    else:
        # 8 lanes per chain need a row stride of 64: D = 40, one component
        model, log_Z, _, _ = gmm_problem(40, seed=7, K=1)
        ns = J.NestedSampler(model, num_live_points=300, s=2, max_samples=300 * 300, gradient_guided=True,
                             lanes_per_chain=lpc)
    reason, state = ns(5)
    b = state.batch
    assert not b.external and int(b.runs_host[0].sampler) == 2
    res = ns.to_results(reason, state)
    ev = np.asarray(res.num_likelihood_evaluations_per_sample)
    say(f"  log Z {res.log_Z_mean:.4f} +- {res.log_Z_uncert:.4f}, analytic {log_Z:.4f}; cap hits "
        f"{int(b.cap_hits().sum())}; min evals of slice samples {ev[ev > 1].min()}")
    assert int(b.cap_hits().sum()) == 0
    assert abs(res.log_Z_mean - log_Z) < 4 * res.log_Z_uncert + 0.15
    assert (ev[ev > 1] >= 2 * ns.num_slices).all()


def test_gradient_guided_fused_and_stepwise_agree():
    """The fused gradient-guided form (device gradient) and the stepwise form (torch autograd of the same
    likelihood as a callable) over 24 seeds each of one MVN problem: both ensembles' mean log Z agree
    with the analytic value and with each other within 4 combined standard errors, and their mean
    evaluation counts per run within 5 %."""
    import torch
    from jaxns_amd.engine import DeviceBatch, RunSpec
    say("gradient-guided fused vs stepwise")
    D = 4
    model, log_Z, mu, Sig = mvn_problem(D, seed=11)
    P = torch.linalg.inv(torch.from_numpy(Sig))
    mu_t = torch.from_numpy(mu)
    c = -0.5 * D * math.log(2 * math.pi) - 0.5 * float(np.linalg.slogdet(Sig)[1])

    def log_likelihood(x):
        r = x - mu_t.to(x.device)
        return c - 0.5 * ((r @ P.to(x.device)) * r).sum(-1)

    def prior_model():
        x = yield J.Prior(J.distributions.Normal(loc=np.zeros(D), scale=np.ones(D)), name="x")
        return x

    callable_model = J.Model(prior_model, log_likelihood)
    R = 24
    out = {}
    for name, m in (("fused", model), ("stepwise", callable_model)):
        specs = [RunSpec(model=m, num_live_points=100, num_slices=2 * D, max_samples=100 * 100, seed=300 + i)
                 for i in range(R)]
        b = DeviceBatch(specs, lanes_per_chain=2, waves_per_block=2, gradient_guided=True)
        assert b.external == (name == "stepwise")
        b.run(dict(dlogZ=math.log(1.0 + 1e-2), max_samples=1e18))
        z = b.log_Z()[:, 0]
        evals = np.array([int(st.num_likelihood_evaluations) for st in b.states()], np.float64)
        out[name] = (z, evals)
        say(f"  {name}: mean log Z {z.mean():.4f} +- {z.std(ddof=1) / math.sqrt(R):.4f} (analytic {log_Z:.4f}); "
            f"mean evaluations {evals.mean():.0f}")
    (zf, ef), (zs, es) = out["fused"], out["stepwise"]
    sf, ss = zf.std(ddof=1) / math.sqrt(R), zs.std(ddof=1) / math.sqrt(R)
    for z, se in ((zf, sf), (zs, ss)):
        assert abs(z.mean() - log_Z) < 4 * se + 0.05
    assert abs(zf.mean() - zs.mean()) < 4 * math.sqrt(sf ** 2 + ss ** 2) + 0.02
    assert abs(ef.mean() / es.mean() - 1) < 0.05


def _one_seed_live_points(spec, D, rng, n=50):
    """n uniform live points and a contour between the best two: every chain starts from the best one."""
    U = rng.uniform(0.02, 0.98, size=(n, D)).astype(np.float32).astype(np.float64)
    L = spec.log_likelihood(U)
    order = np.argsort(L)
    return U, L, U[order[-1]], float(L[order[-2]])


def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


@pytest.mark.parametrize("fam,D,lpc", [("mvn", 6, 2), ("mvn", 20, 4), ("gmm", 5, 2), ("gmm", 40, 8)])
def test_gradient_slice_moves_up_the_gradient(fam, D, lpc):
    """UniDimSliceSampler(gradient_slice=True) (uni_slice_sampler.py:202-214): every slice runs along the
    unit gradient at its seed over [0, right].  From a single seed point, with two slices and one phantom
    per sample, the phantom (the point after the first slice) lies on the ray from the seed along the
    central-difference gradient of the oracle's FP64 likelihood, and the sample on the ray from the phantom
    along the gradient there: 1 - cos < 1e-6 and a positive step.  Each sample spends one evaluation per
    slice on the gradient (at least 4 in all); with gradient_guided as well, the samples are the same bit
    for bit and every sample counts one more evaluation per slice (the reference's second grad_fn call at
    each accepted point)."""
    from jaxns_amd.samplers import EphemeralState, LivePointCollection, UniDimSliceSampler
    say(f"gradient slice {fam} D={D} lpc={lpc}")
    rng = np.random.default_rng(100 + D)
    if fam == "mvn":
        model, _, mu, Sig = mvn_problem(D, seed=D)
        spec = mvn_oracle_spec(D, mu, Sig)
    else:
        model, _, means, sig = gmm_problem(D, seed=D)  # D = 40: row stride 64, 8 lanes per chain
        spec = gmm_oracle_spec(D, means, sig)
    U, L, U0, contour = _one_seed_live_points(spec, D, rng)
    lp = LivePointCollection(sender_node_idx=np.zeros(len(L), np.int64), U_sample=U, log_L_constraint=np.zeros(len(L)),
                             log_L=L, num_likelihood_evaluations=np.ones(len(L), np.int64))
    n = 256
    out = {}
    for gg in (False, True):
        sampler = UniDimSliceSampler(model, num_slices=2, num_phantom_save=1, midpoint_shrink=True, gradient_slice=True,
                                     gradient_guided=gg, lanes_per_chain=lpc)
        assert sampler._sampler_code == (4 if gg else 3)
        state = sampler.pre_process(EphemeralState(key=0, live_points_collection=lp))
        out[gg] = sampler.get_samples((41, 3), contour, state, n)
    (s, ph), (s2, ph2) = out[False], out[True]
    assert np.all(s.log_L > contour) and np.all(ph.log_L > contour)
    assert np.all((s.U_sample >= 0) & (s.U_sample <= 1)) and np.all((ph.U_sample >= 0) & (ph.U_sample <= 1))
    g0 = _unit(_fd_grad(spec, U0[None, :])[0])
    gp = _unit(_fd_grad(spec, ph.U_sample))
    d1 = ph.U_sample - U0
    d2 = s.U_sample - ph.U_sample
    n1, n2 = np.linalg.norm(d1, axis=1), np.linalg.norm(d2, axis=1)
    m1, m2 = n1 > 1e-3, n2 > 1e-3  # steps long enough that FP32 rounding of the points leaves the angle
    say(f"  steps > 1e-3: {m1.sum()} / {n}, {m2.sum()} / {n}; evals {s.num_likelihood_evaluations.min()}.."
        f"{s.num_likelihood_evaluations.max()}")
    assert m1.sum() > n // 2 and m2.sum() > n // 2
    cos1 = (d1[m1] @ g0) / n1[m1]
    cos2 = np.sum(d2[m2] * gp[m2], axis=1) / n2[m2]
    say(f"  1 - cos: first slice {np.max(1 - cos1):.2e}, second {np.max(1 - cos2):.2e}")
    assert np.all(1 - cos1 < 1e-6) and np.all(1 - cos2 < 1e-6)
    assert np.all(s.num_likelihood_evaluations >= 4)
    # gradient_guided on top: the same chains, one more evaluation per slice
    assert np.array_equal(s2.U_sample, s.U_sample) and np.array_equal(s2.log_L, s.log_L)
    assert np.array_equal(ph2.U_sample, ph.U_sample)
    assert np.all(s2.num_likelihood_evaluations == s.num_likelihood_evaluations + 2)


def test_gradient_slice_whole_runs():
    """Whole runs with gradient_slice (DeviceBatch(gradient_slice=True), sampler code 3) and with
    gradient_slice and gradient_guided (code 4) on four seeds of an MVN problem: both terminate with no
    step-cap hits and collect the same samples bit for bit; per sample, the evaluation counts are equal
    for the initial points and the phantoms and one per slice apart for every slice sample; the best sample
    ends within 0.5 nat of the likelihood's peak (the sampler is meant to climb)."""
    from jaxns_amd.engine import DeviceBatch, RunSpec
    say("gradient slice whole runs")
    D, N = 6, 200
    model, _, mu, Sig = mvn_problem(D, seed=5)
    ns = 2 * D
    specs = [RunSpec(model=model, num_live_points=N, num_slices=ns, k_phantom=1, max_samples=N * 150, seed=900 + i)
             for i in range(4)]
    res = {}
    for gg in (False, True):
        b = DeviceBatch(specs, lanes_per_chain=2, waves_per_block=2, gradient_slice=True, gradient_guided=gg)
        assert not b.external and int(b.runs_host[0].sampler) == (4 if gg else 3)
        b.run(dict(dlogZ=math.log(1.0 + 1e-2), max_samples=1e18), init_efficiency_threshold=0.0)
        assert int(b.cap_hits().sum()) == 0
        res[gg] = (b, [b.sample_collection(r) for r in range(len(specs))], b.states())
    (b, cols, sts), (_, cols2, sts2) = res[False], res[True]
    peak = float(model.log_likelihood.log_const)  # log L at x = mu
    for r in range(len(specs)):
        c, c2 = cols[r], cols2[r]
        m = int(sts[r].num_samples)
        assert int(sts2[r].num_samples) == m
        assert np.array_equal(c.U_samples[:m], c2.U_samples[:m]) and np.array_equal(c.log_L[:m], c2.log_L[:m])
        diff = c2.num_likelihood_evaluations[:m].astype(np.int64) - c.num_likelihood_evaluations[:m]
        assert set(np.unique(diff)) <= {0, ns}
        # the collection holds the dead points of every shell, their phantoms and finally the live set:
        # each initial point once (one evaluation: the MVN log L is finite everywhere), every other
        # non-phantom entry a slice sample (ns gradient steps and at least ns proposals)
        ph = c.phantom[:m]
        ev = c.num_likelihood_evaluations[:m]
        init = ~ph & (ev == 1)
        slice_samples = ~ph & ~init
        assert init.sum() == N and np.all(ev[slice_samples] >= 2 * ns)
        assert np.all(diff[ph] == 0) and np.all(diff[init] == 0) and np.all(diff[slice_samples] == ns)
        say(f"  run {r}: {m} samples, best log L {c.log_L[:m].max():.4f}, peak {peak:.4f}")
        assert c.log_L[:m].max() > peak - 0.5
