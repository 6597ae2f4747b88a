"""GPU parity of _to_results (nested_samplers/sharded/sharded_static.py:652-773): the device's results
of finished runs against the oracle's to_results on the device's own sample collection, the per-sample
finals (normalised log dp, H, transform, log posterior density) against FP64 numpy and scipy, and the
untrimmed form (trim=False) against the oracle's dynamic count and evidence scan."""
import numpy as np
import pytest
from scipy import stats
from scipy.special import ndtri

import jaxns_amd as J
import oracle as O
from jaxns_amd import distributions as tfpd

from problems import mvn_problem

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _collection(b, r):
    """The run's sample collection as host arrays (its first min(num_samples, max_samples) entries)."""
    st = b.states()[r]
    d = b.runs_host[r]
    M = int(min(st.num_samples, d.max_samples))
    o = int(d.samp_off)
    return (b.samp_logL[o:o + M].cpu().numpy(), b.samp_sender[o:o + M].cpu().numpy().astype(np.int64),
            b.samp_evals[o:o + M].cpu().numpy().astype(np.int64), b.samp_phantom[o:o + M].cpu().numpy())


def _tril_model(D=5, seed=3):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(D, D))
    L = np.linalg.cholesky(A @ A.T / D + 0.5 * np.eye(D))
    loc = rng.normal(size=D)
    mu = loc + 0.3 * rng.normal(size=D)

    def prior_model():
        x = yield J.Prior(tfpd.MultivariateNormalTriL(loc=loc, scale_tril=L), name="x")
        return x

    return J.Model(prior_model, J.MultivariateNormalLikelihood(mu, cov=0.2 * np.eye(D)))


def _scalar_model():
    def prior_model():
        a = yield J.Prior(tfpd.Uniform(-2.0, 3.0), name="a")
        b = yield J.Prior(tfpd.Normal(np.zeros(3), 2.0 * np.ones(3)), name="b")
        return a, b

    means = np.array([[0.5, 0.1, -0.2, 0.3]])
    return J.Model(prior_model, J.GaussianMixtureLikelihood(means, np.array([0.4])))


@pytest.mark.parametrize("which,k", [("mvn", 0), ("mvn", 2), ("tril", 0), ("scalar", 1)])
def test_results_match_oracle_to_results(which, k):
    """log Z, its uncertainty, ESS, H, sample order and live-point counts of the device's results equal
    the oracle's to_results over the same sample collection (1e-12 / exact); log dp is the normalised
    log dZ of the oracle's evidence scan; samples = Model.transform(U) and log_posterior_density =
    log L + log prior (scipy's logpdf for the MVN-TriL prior) to 1e-12."""
    say(f"results vs oracle: {which} k={k}")
    if which == "mvn":
        model, _, _, _ = mvn_problem(6, seed=2)
    elif which == "tril":
        model = _tril_model()
    else:
        model = _scalar_model()
    ns = J.NestedSampler(model, num_live_points=150, s=3, k=k, max_samples=150 * 60)
    reason, state = ns(7)
    res = ns.to_results(reason, state)
    b = state.batch
    sL, sS, sE, sP = _collection(b, 0)
    o = O.to_results(sL, sS, sE, sP)
    M = sL.size
    assert res.total_num_samples == M == o["num_samples"]
    assert np.array_equal(res.num_live_points_per_sample, o["num_live_points"])
    assert np.array_equal(res.log_L_samples, sL[o["samples_indices"]])
    assert res.total_num_likelihood_evaluations == o["total_likelihood_evaluations"]
    assert res.total_phantom_samples == o["total_phantom_samples"] and (k == 0) == (res.total_phantom_samples == 0)
    for f in ("log_Z_mean", "log_Z_uncert", "ESS", "H_mean", "log_efficiency"):
        np.testing.assert_allclose(getattr(res, f), o[f], rtol=1e-11, atol=1e-11, err_msg=f)
    per = o["per_sample"]
    np.testing.assert_allclose(res.log_X_mean, per[:, 1], rtol=1e-11, atol=1e-11)
    ldz = per[:, 6]
    mx = ldz.max()
    want_dp = ldz - (mx + np.log(np.sum(np.exp(ldz - mx))))
    fin = np.isfinite(want_dp)
    assert np.array_equal(np.isfinite(res.log_dp_mean), fin)
    np.testing.assert_allclose(res.log_dp_mean[fin], want_dp[fin], rtol=1e-11, atol=1e-11)
    U = res.U_samples
    assert U.dtype == np.float64 and U.shape == (M, model.U_ndims)
    want_X = model.transform(U)
    for name, x in want_X.items():
        got = res.samples[name]
        np.testing.assert_allclose(np.reshape(got, np.shape(x)), x, rtol=1e-12, atol=1e-12, err_msg=name)
    lp = model.log_prob_prior_numpy(U)
    if which == "tril":
        d = model.priors[0].dist
        lp_scipy = stats.multivariate_normal(d.loc, d.scale_tril @ d.scale_tril.T).logpdf(want_X["x"])
        np.testing.assert_allclose(lp, lp_scipy, rtol=1e-10)
    np.testing.assert_allclose(res.log_posterior_density, res.log_L_samples + lp, rtol=1e-12, atol=1e-12)
    if which == "scalar":
        assert res.samples["a"].shape == (M,) and res.samples["b"].shape == (M, 3)


def test_untrimmed_results():
    """to_results(trim=False) (public.py:177-196, sharded_static.py:652-705): every row of the sample
    collection, the tree count and evidence scan stopping at num_samples (the oracle's dynamic forms);
    its first num_samples rows equal the trimmed results, the placeholders past them have log L = +inf,
    U = 0.5 and no weight, and trim_results cuts it back to the trimmed results."""
    say("untrimmed results")
    model, _, _, _ = mvn_problem(4, seed=5)
    ns = J.NestedSampler(model, num_live_points=100, s=3, max_samples=100 * 80)
    reason, state = ns(3)
    a = ns.to_results(reason, state)
    u = ns.to_results(reason, state, trim=False)
    b = state.batch
    ms = int(b.runs_host[0].max_samples)
    M = a.total_num_samples
    assert M < ms and u.total_num_samples == M and u.log_L_samples.shape == (ms,)
    for f in ("log_Z_mean", "log_Z_uncert", "ESS", "H_mean", "total_num_likelihood_evaluations",
              "total_phantom_samples"):
        np.testing.assert_allclose(getattr(u, f), getattr(a, f), rtol=1e-12, err_msg=f)
    o = int(b.runs_host[0].samp_off)
    fullL = b.samp_logL[o:o + ms].cpu().numpy()
    fullS = b.samp_sender[o:o + ms].cpu().numpy().astype(np.int64)
    si, nl = O.count_crossed_edges(fullS, fullL, num_samples=M)
    assert np.array_equal(u.num_live_points_per_sample, nl)
    assert np.array_equal(u.log_L_samples, fullL[si])
    fin, per = O.compute_evidence_stats(fullL[si], nl.astype(np.float64), stop_idx=M)
    np.testing.assert_allclose(u.log_X_mean, per[:, 1], rtol=1e-11, atol=1e-11)
    assert np.all(np.isposinf(u.log_L_samples[M:]))
    assert np.all(u.U_samples[M:] == 0.5)
    assert np.all(np.isneginf(u.log_dp_mean[M:]))
    t = ns.trim_results(u)
    for f in ("log_L_samples", "U_samples", "log_dp_mean", "log_X_mean", "num_live_points_per_sample",
              "log_posterior_density"):
        np.testing.assert_array_equal(getattr(t, f), getattr(a, f), err_msg=f)


def test_results_finals_large_vs_numpy():
    """hipns_results_finals over 300 001 synthetic samples (147 tiles): normalised log dp and both forms
    of H against FP64 numpy, with -inf weights at the start as the scan produces them."""
    import ctypes as C

    import torch
    from jaxns_amd import _lib
    from jaxns_amd._lib import check, lib, ptr
    say("results finals 300001")
    rng = np.random.default_rng(0)
    M = 300_001
    ldz = rng.normal(size=M) * 5 - 40
    ldz[:17] = -np.inf
    lL = np.sort(rng.normal(size=M))
    lX = -np.sort(rng.exponential(size=M)).cumsum() / M
    log_Z = 1.25
    dz, dl, dx = (torch.from_numpy(v).cuda() for v in (ldz, lL, lX))
    out = torch.empty(M, dtype=torch.float64, device="cuda")
    H = torch.empty(4, dtype=torch.float64, device="cuda")
    ws = torch.empty(int(lib().hipns_results_finals_workspace_size(M)), dtype=torch.uint8, device="cuda")
    check(lib().hipns_results_finals(ptr(dz), ptr(dl), ptr(dx), M, log_Z, ptr(out), ptr(H), ptr(ws), ws.numel(),
                                     _lib.stream_ptr()))
    got, Hh = out.cpu().numpy(), H.cpu().numpy()
    mx = ldz.max()
    ln = mx + np.log(np.sum(np.exp(ldz - mx)))
    want = ldz - ln
    assert np.array_equal(np.isneginf(got), np.isneginf(want))
    np.testing.assert_allclose(got[17:], want[17:], rtol=1e-13)
    p = np.exp(want)
    np.testing.assert_allclose(Hh[0], -(np.sum(p[17:] * lL[17:]) - log_Z), rtol=1e-12)
    np.testing.assert_allclose(Hh[1], np.sum(p * lX), rtol=1e-12)
    assert Hh[2] == Hh[0]
    np.testing.assert_allclose(Hh[3], ln, rtol=1e-14)
