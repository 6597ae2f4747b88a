"""Pins the CPU oracle (tests only) against the reference's known answers and the mathematics.

Golden vectors are the reference's own unit-test values (jaxns 2.6.9):
  internals/tests/test_tree_structure.py:19-80, test_cumulative_ops.py:95-126,
  test_log_semiring.py:8-67, test_stats.py:7-16,
  nested_samplers/sharded/tests/test_sharded_static.py:4-7.
Everything else is pinned by closed forms, brute force or library routines.
"""
import math

import numpy as np
import pytest
from scipy import stats
from scipy.special import ndtri

import oracle as O

from problems import linreg_oracle_spec, linreg_problem, mvn_oracle_spec


# ---------------------------------------------------------------- tree counts
def test_tree_kat_basic():
    si, nl = O.count_crossed_edges([0, 0, 0, 1, 2, 3], [1, 2, 3, 4, 5, 6])
    assert list(si) == [0, 1, 2, 3, 4, 5]
    assert list(nl) == [3, 3, 3, 3, 2, 1]
    si2, nl2 = O.count_intervals_naive([0, 0, 0, 1, 2, 3], [1, 2, 3, 4, 5, 6])
    assert list(si2) == list(si) and list(nl2) == list(nl)


def test_tree_kat_permuted():
    a = O.count_crossed_edges([0, 0, 0, 1, 3, 2], [1, 2, 3, 4, 6, 5])
    b = O.count_intervals_naive([0, 0, 0, 1, 3, 2], [1, 2, 3, 4, 6, 5])
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_tree_kat_num_samples():
    si1, nl1 = O.count_crossed_edges([0, 0, 0, 1, 2, 3, 4, 5, 0, 0], [1, 2, 3, 4, 5, 6, 7, 8, np.inf, np.inf], 8)
    si2, nl2 = O.count_crossed_edges([0, 0, 0, 1, 2, 3, 4, 5], [1, 2, 3, 4, 5, 6, 7, 8])
    assert np.array_equal(si1[:8], si2) and np.array_equal(nl1[:8], nl2)
    assert np.all(nl1[8:] == 0)


@pytest.mark.parametrize("seed", [42, 1, 2])
def test_tree_random_vs_naive(seed):
    # the random-tree construction of test_tree_structure.py:73-92
    rng = np.random.RandomState(seed)
    log_L = [0.0]
    parent = []
    for idx in range(10):
        log_L.append(log_L[idx] + rng.uniform(low=0, high=1 - log_L[idx]) ** 4)
        parent.append(idx)
    for idx in range(10):
        for _ in range(5):
            log_L.append(rng.uniform(low=log_L[idx], high=1.0))
            parent.append(idx)
    a = O.count_crossed_edges(parent, log_L[1:])
    b = O.count_intervals_naive(parent, log_L[1:])
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


# ---------------------------------------------------------------- log space
def test_log_semiring_kats():
    for a, b in [(5.0, 10.0), (0.0, 10.0)]:
        la, lb = np.log(a) if a > 0 else -np.inf, np.log(b)
        assert math.isclose(math.exp(O.logaddexp(la, lb)), a + b, rel_tol=1e-12)
        v, s = O.signed_logaddexp(la, 1.0, lb, -1.0)
        assert math.isclose(s * math.exp(v), a - b, rel_tol=1e-12)
    v, s = O.signed_logaddexp(np.log(5.0), -1.0, np.log(10.0), 1.0)
    assert math.isclose(s * math.exp(v), 5.0, rel_tol=1e-12)
    assert O.logaddexp(-np.inf, -np.inf) == -np.inf
    assert O.logaddexp(np.inf, np.inf) == np.inf
    rng = np.random.default_rng(0)
    x = rng.normal(size=(100, 2)) * 30
    for a, b in x:
        assert math.isclose(O.logaddexp(a, b), np.logaddexp(a, b), rel_tol=1e-14, abs_tol=1e-14)
    u = rng.normal(size=50)
    np.testing.assert_allclose(O.cumulative_logsumexp(u), np.logaddexp.accumulate(u), rtol=1e-13)


def test_linear_to_log_stats_lognormal():
    # test_stats.py:7-16: Z = exp(N(0,1)) -> E[log Z] ~ 0, Var[log Z] ~ 1
    rng = np.random.default_rng(42)
    Z = np.exp(rng.normal(size=1_000_000))
    mu, var = O.linear_to_log_stats(np.log(Z.mean()), np.log(Z.var() + Z.mean() ** 2))
    assert abs(mu) < 1e-2 and abs(var - 1) < 1e-2


# ---------------------------------------------------------------- evidence recurrence
def test_evidence_closed_forms():
    # X_i = prod n/(n+1), X2_i = prod n/(n+2); constant L = 1 gives Z_M = (1 - X_1)/2 + X_1 - X_M
    rng = np.random.default_rng(1)
    n = rng.integers(1, 40, size=300).astype(float)
    fin, per = O.compute_evidence_stats(np.zeros(300), n)
    X = np.cumprod(n / (n + 1))
    X2 = np.cumprod(n / (n + 2))
    np.testing.assert_allclose(per[:, 1], np.log(X), rtol=1e-12)
    np.testing.assert_allclose(per[:, 2], np.log(X2), rtol=1e-12)
    Zc = 0.5 * (1 - X[0]) + X[0] - X
    np.testing.assert_allclose(per[:, 3], np.log(Zc), rtol=1e-12)
    assert fin[0] == 0.0


def test_evidence_moments_monte_carlo():
    """E[Z], E[Z^2], E[ZX], E[X^2] of the shrinkage process T_i ~ Beta(n_i, 1),
    Z = sum (X_{i-1} - X_i) (L_{i-1} + L_i)/2, simulated directly."""
    rng = np.random.default_rng(7)
    M = 12
    n = rng.integers(2, 8, size=M).astype(float)
    L = np.sort(rng.uniform(0.2, 2.0, size=M))
    fin, per = O.compute_evidence_stats(np.log(L), n)
    S = 400_000
    T = rng.uniform(size=(S, M)) ** (1.0 / n)
    X = np.cumprod(T, axis=1)
    Xp = np.concatenate([np.ones((S, 1)), X[:, :-1]], axis=1)
    m = 0.5 * (np.concatenate([[0.0], L[:-1]]) + L)
    Z = np.sum((Xp - X) * m, axis=1)
    ZX = Z * X[:, -1]
    for got, want in [(fin[3], Z.mean()), (fin[5], (Z ** 2).mean()), (fin[4], ZX.mean()),
                      (fin[2], (X[:, -1] ** 2).mean()), (fin[1], X[:, -1].mean())]:
        assert abs(math.exp(got) / want - 1) < 1e-2, (got, want)
    # dZ2 = sum E[dZ^2]
    dZ = (Xp - X) * m
    assert abs(math.exp(fin[7]) / (dZ ** 2).sum(axis=1).mean() - 1) < 1e-2


def test_evidence_dynamic_stop():
    # cumulative_op_dynamic semantics (test_cumulative_ops.py:95-126): outputs past stop hold the init value
    L = np.log([1.0, 2.0, 3.0])
    n = np.array([3.0, 2.0, 1.0])
    fin, per = O.compute_evidence_stats(L, n, 2)
    fin3, per3 = O.compute_evidence_stats(L, n)
    np.testing.assert_array_equal(per[:2], per3[:2])
    np.testing.assert_array_equal(per[2], O.evidence_init())
    np.testing.assert_array_equal(fin, per3[1])


# ---------------------------------------------------------------- special functions / rng
def test_ndtri_vs_scipy():
    p = np.concatenate([np.logspace(-300, -1, 300), np.linspace(0.01, 0.99, 999), 1 - np.logspace(-16, -1, 100)])
    err = np.abs(O.ndtri(p) - ndtri(p)) / np.maximum(1, np.abs(ndtri(p)))
    assert err.max() < 1e-14


def test_philox_kat():
    assert [int(x) for x in O.philox4x32([0, 0, 0, 0], [0, 0])] == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert [int(x) for x in O.philox4x32([0xffffffff] * 4, [0xffffffff] * 2)] == [0x408f276d, 0x41c83b0e,
                                                                                   0xa20bc7c6, 0x6d5451fd]
    assert [int(x) for x in O.philox4x32([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                                         [0xa4093822, 0x299f31d0])] == [0xd16cfe09, 0x94fdcceb, 0x5001e420,
                                                                        0x24126ea1]


def test_seed_index_brute_force():
    rng = np.random.default_rng(3)
    for _ in range(200):
        N = int(rng.integers(2, 50))
        L = np.sort(np.round(rng.normal(size=N), 1))
        contour = L[int(rng.integers(0, N))]
        u = float(rng.uniform())
        sel = L > contour
        M = sel.sum()
        cnt = np.cumsum(sel)
        want = int(np.argmax(cnt >= M * (1 - u))) if M > 0 else 0
        assert O.seed_index(L, contour, u) == want
        if M > 0:  # the integer form the device uses
            a = int(np.argmax(sel))
            assert want == a + math.ceil(M * (1 - u)) - 1


# ---------------------------------------------------------------- slice sampler
def test_slice_bounds_on_cube_boundary():
    rng = np.random.default_rng(5)
    for dt in (np.float64, np.float32):
        for _ in range(500):
            D = int(rng.integers(1, 12))
            U0 = rng.uniform(size=D)
            n = rng.normal(size=D)
            n /= np.linalg.norm(n)
            lft, rgt = O.slice_bounds(U0, n, dtype=dt)
            assert lft <= 0 <= rgt
            tol = 1e-12 if dt == np.float64 else 1e-5
            for t in (lft, rgt):
                x = U0 + t * n
                assert np.all(x >= -tol) and np.all(x <= 1 + tol)
                assert min(np.min(np.abs(x)), np.min(np.abs(1 - x))) < tol


def _ball_spec(D, mu):
    return O.ModelSpec(O.FAM_MVN, D, O.PRIOR_UNIFORM, np.zeros(D), np.ones(D), mvn_mu=mu,
                       mvn_chol=0.2 * np.eye(D))


def test_slice_chain_uniform_in_ball_ks():
    """Constraint logL > c of an isotropic Gaussian is a ball; slice chains started at its
    centre and run s*D slices must give radius^D ~ U[0, r^D] (KS)."""
    D = 3
    mu = np.full(D, 0.5)
    spec = _ball_spec(D, mu)
    r = 0.3
    contour = float(spec.log_likelihood(mu + np.array([r, 0, 0]))[0])
    rng = np.random.default_rng(11)
    ns = 15
    radii = []
    for c in range(1500):
        z = rng.normal(size=(ns, D))
        dirs = z / np.linalg.norm(z, axis=1, keepdims=True)
        tu = rng.uniform(size=(ns, 200))
        o = O.slice_chain(spec, mu, float(spec.log_likelihood(mu)[0]), contour, ns, dirs, tu)
        radii.append(np.linalg.norm(o["U"] - mu))
        assert o["logL"] > contour
    v = (np.array(radii) / r) ** D
    assert stats.kstest(v, "uniform").pvalue > 1e-3


def test_slice_chain_plain_shrink_1d_uniform():
    """midpoint_shrink=False is standard shrink-to-rejected slice sampling; with a box constraint
    in D=1 the chain output is uniform on the accepted interval (KS)."""
    D = 1
    mu = np.array([0.6])
    spec = _ball_spec(D, mu)
    contour = float(spec.log_likelihood(np.array([0.6 + 0.25]))[0])  # accepted set (0.35, 0.85)
    rng = np.random.default_rng(2)
    out = []
    for c in range(2000):
        o = O.slice_chain(spec, mu, float(spec.log_likelihood(mu)[0]), contour, 3, np.ones((3, 1)),
                          rng.uniform(size=(3, 100)), midpoint_shrink=False)
        out.append(o["U"][0])
    v = (np.array(out) - 0.35) / 0.5
    assert stats.kstest(v, "uniform").pvalue > 1e-3


def test_slice_chain_phantoms_and_evals():
    D = 4
    rng = np.random.default_rng(9)
    mu = np.full(D, 0.5)
    spec = _ball_spec(D, mu)
    contour = float(spec.log_likelihood(mu + 0.3)[0])
    ns, k = 8, 3
    z = rng.normal(size=(ns, D))
    dirs = z / np.linalg.norm(z, axis=1, keepdims=True)
    tu = rng.uniform(size=(ns, 100))
    # midpoint_shrink=False: no alpha schedule, so a chain of ns-1 slices repeats the first ns-1 slices
    o = O.slice_chain(spec, mu, float(spec.log_likelihood(mu)[0]), contour, ns, dirs, tu, k_phantom=k,
                      midpoint_shrink=False)
    assert o["evals"] == int(o["steps"].sum())
    # phantoms are the accepted points of slices ns-k-1 .. ns-2; the last slice's point is the output
    o_short = O.slice_chain(spec, mu, float(spec.log_likelihood(mu)[0]), contour, ns - 1, dirs[:-1], tu[:-1],
                            midpoint_shrink=False)
    np.testing.assert_array_equal(o["phantom_U"][-1], o_short["U"])
    assert np.all(o["phantom_logL"] > contour)


# ---------------------------------------------------------------- termination
def test_termination_bits():
    r = O.init_register()
    done, why = O.determine_termination(dict(efficiency_threshold=0.1, dlogZ=0.0, max_samples=100), r)
    assert done and why == 1 << 6  # the uniform phase stops before its first shell
    done, why = O.determine_termination(dict(dlogZ=np.log(1 + 1e-3), max_samples=100), r)
    assert not done and why == 0
    r.num_samples_used = 100
    r.plateau = 1
    done, why = O.determine_termination(dict(max_samples=100), r)
    assert done and why == (1 << 0) + (1 << 7)


# ---------------------------------------------------------------- full run
def test_oracle_full_run_analytic():
    from problems import mvn_problem
    _, log_Z, mu, Sig = mvn_problem(6, seed=4)
    spec = mvn_oracle_spec(6, mu, Sig)
    zs = []
    for seed in range(4):
        r = O.ns_run(spec, num_live_points=200, num_slices=30, max_samples=200 * 100, seed=seed, num_threads=4)
        zs.append((r["log_Z_mean"], r["log_Z_uncert"]))
        assert r["termination_reason"] & 4
    zs = np.array(zs)
    assert abs(zs[:, 0].mean() - log_Z) < 3 * zs[:, 1].mean() / 2 + 0.05


def test_oracle_full_run_device_direction_form():
    """The run option dir16 (slice directions from 16-bit uniform pairs, eight dimensions per Philox
    block, the form the device used to draw) changes the run's random stream, and its log Z agrees with
    the analytic value as the FP64 Box-Muller form's does."""
    from problems import mvn_problem
    _, log_Z, mu, Sig = mvn_problem(6, seed=4)
    spec = mvn_oracle_spec(6, mu, Sig)
    zs = []
    for seed in range(4):
        a = O.ns_run(spec, num_live_points=200, num_slices=30, max_samples=200 * 100, seed=seed, num_threads=4)
        b = O.ns_run(spec, num_live_points=200, num_slices=30, max_samples=200 * 100, seed=seed, num_threads=4,
                     dir16=True)
        assert a["total_likelihood_evaluations"] != b["total_likelihood_evaluations"] or \
            not np.array_equal(a["samples_logL"], b["samples_logL"])
        assert b["termination_reason"] & 4
        zs.append((b["log_Z_mean"], b["log_Z_uncert"]))
    zs = np.array(zs)
    assert abs(zs[:, 0].mean() - log_Z) < 3 * zs[:, 1].mean() / 2 + 0.05


def test_linreg_loglik_vs_scipy():
    """Oracle regression likelihood at x(u) = tau ndtri(u) vs scipy's normal logpdf summed over rows."""
    _, _, lik = linreg_problem(200, 7, seed=3, sigma=0.5, tau=2.0)
    spec = linreg_oracle_spec(lik, 2.0)
    rng = np.random.default_rng(0)
    U = rng.uniform(0.01, 0.99, size=(10, 7))
    got = spec.log_likelihood(U)
    W = 2.0 * ndtri(U)
    want = np.array([stats.norm.logpdf(lik.y, lik.X @ w, 0.5).sum() for w in W])
    np.testing.assert_allclose(got, want, rtol=1e-12)


# ---------------------------------------------------------------- ESS, prior density
def test_ess_kish():
    """effective_sample_size_kish (stats.py:77-86) = (sum dZ)^2 / sum dZ^2: M for equal weights, and the
    ratio of the direct sums for random ones."""
    for M in (1, 7, 1000):
        w = np.full(M, 0.37)
        assert O.ess_kish(np.log(w.sum()), np.log((w * w).sum())) == pytest.approx(M, rel=1e-12)
    rng = np.random.default_rng(3)
    w = rng.exponential(size=500) ** 3
    want = w.sum() ** 2 / (w * w).sum()
    assert O.ess_kish(np.log(w.sum()), np.log((w * w).sum())) == pytest.approx(want, rel=1e-12)


def test_log_prob_prior_vs_scipy():
    """Model.log_prob_prior (framework/model.py:178-185): the prior log density at X(U), against scipy's
    logpdf of the Normal, Uniform and lower-triangular multivariate normal priors."""
    rng = np.random.default_rng(5)
    D = 4
    U = rng.uniform(0.02, 0.98, size=(20, D))
    loc, scale = rng.normal(size=D), rng.uniform(0.5, 2.0, size=D)
    spec = O.ModelSpec(O.FAM_MVN, D, O.PRIOR_NORMAL, loc, scale, mvn_mu=np.zeros(D), mvn_chol=np.eye(D))
    X = loc + scale * ndtri(U)
    np.testing.assert_allclose(spec.log_prob_prior(U), stats.norm.logpdf(X, loc, scale).sum(1), rtol=1e-12)
    low, width = rng.normal(size=D), rng.uniform(0.5, 3.0, size=D)
    spec = O.ModelSpec(O.FAM_MVN, D, O.PRIOR_UNIFORM, low, width, mvn_mu=np.zeros(D), mvn_chol=np.eye(D))
    want = stats.uniform.logpdf(low + width * U, low, width).sum(1)
    np.testing.assert_allclose(spec.log_prob_prior(U), want, rtol=1e-12)
    A = rng.normal(size=(D, D))
    L = np.linalg.cholesky(A @ A.T + D * np.eye(D))
    spec = O.ModelSpec(O.FAM_MVN, D, O.PRIOR_MVN, loc, np.ones(D), prior_tril=L, mvn_mu=np.zeros(D),
                       mvn_chol=np.eye(D))
    X = loc + ndtri(U) @ L.T
    want = stats.multivariate_normal(loc, L @ L.T).logpdf(X)
    np.testing.assert_allclose(spec.log_prob_prior(U), want, rtol=1e-12)


# ---------------------------------------------------------------- slice rules written out by hand
def _unit_spec():
    return O.ModelSpec(O.FAM_MVN, 1, O.PRIOR_NORMAL, np.zeros(1), np.ones(1), mvn_mu=np.zeros(1),
                       mvn_chol=np.eye(1))


def test_midpoint_shrink_known_answer():
    """One 1-D slice from U0 = 0.5 (direction 1, bracket [-0.5, 0.5]) with replayed likelihood values
    (reject, reject, accept) and picks u = 0.9, 0.1, 0.7, every step written out from
    _pick_point_in_interval / _shrink_interval (uni_slice_sampler.py:67-111) with alpha = 0.5 (one slice:
    linspace(0.5, 1, 1) = [0.5]):
      t1 = -0.5 + 0.9 * 1.0 = 0.4                rejected, t > 0: right = 0.5 * 0.4 = 0.2
      t2 = -0.5 + 0.1 * (0.2 + 0.5) = -0.43      rejected, t < 0: left = 0.5 * -0.43 = -0.215
      t3 = -0.215 + 0.7 * (0.2 + 0.215) = 0.0755 accepted: U = 0.5755
    Without the midpoint rule the bounds move to t itself:
      t2 = -0.5 + 0.1 * 0.9 = -0.41, t3 = -0.41 + 0.7 * (0.4 + 0.41) = 0.157."""
    spec = _unit_spec()
    tu = np.array([[0.9, 0.1, 0.7] + [0.5] * 5])
    rep = np.array([[-np.inf, -np.inf, 1.0] + [1.0] * 5])
    o = O.slice_chain(spec, np.array([0.5]), 0.5, 0.0, 1, np.ones((1, 1)), tu, logL_replay=rep)
    np.testing.assert_allclose(o["trace_t"][0, :3], [0.4, -0.43, 0.0755], rtol=0, atol=1e-15)
    assert o["steps"][0] == 3 and o["evals"] == 3
    assert o["U"][0] == pytest.approx(0.5755, abs=1e-15) and o["logL"] == 1.0
    o = O.slice_chain(spec, np.array([0.5]), 0.5, 0.0, 1, np.ones((1, 1)), tu, logL_replay=rep,
                      midpoint_shrink=False)
    np.testing.assert_allclose(o["trace_t"][0, :3], [0.4, -0.41, 0.157], rtol=0, atol=1e-15)
    # two slices: alpha = linspace(0.5, 1, 2) = [0.5, 1.0]; the second slice starts from the first's
    # point 0.5755 (bracket [-0.5755, 0.4245]) and with alpha = 1 shrinks to the rejected point
    tu2 = np.array([[0.9, 0.1, 0.7] + [0.5] * 5, [0.0, 0.5] + [0.5] * 6])
    rep2 = np.array([[-np.inf, -np.inf, 1.0] + [1.0] * 5, [-np.inf, 2.0] + [2.0] * 6])
    o = O.slice_chain(spec, np.array([0.5]), 0.5, 0.0, 2, np.ones((2, 1)), tu2, logL_replay=rep2)
    np.testing.assert_allclose(o["trace_t"][1, :2], [-0.5755, -0.5755 + 0.5 * (0.4245 + 0.5755)], atol=1e-15)
    assert o["logL"] == 2.0 and o["evals"] == 5


def test_plateau_acceptance_rule():
    """_new_proposal's loop condition (uni_slice_sampler.py:160-166): a proposal is accepted when
    logL > contour, or when the seed and the proposal both sit exactly on the contour (a plateau);
    a proposal on the contour from a seed above it is rejected."""
    spec = _unit_spec()
    tu = np.full((1, 8), 0.25)
    # seed on the contour, proposal on the contour: accepted at once
    o = O.slice_chain(spec, np.array([0.5]), 1.0, 1.0, 1, np.ones((1, 1)), tu, logL_replay=np.full((1, 8), 1.0))
    assert o["steps"][0] == 1 and o["logL"] == 1.0
    # seed above the contour: proposals on the contour are rejected until one is above it
    rep = np.array([[1.0, 1.0, 1.5, 0, 0, 0, 0, 0]])
    o = O.slice_chain(spec, np.array([0.5]), 2.0, 1.0, 1, np.ones((1, 1)), tu, logL_replay=rep)
    assert o["steps"][0] == 3 and o["logL"] == 1.5
    # seed on the contour, proposal below it: rejected
    rep = np.array([[0.5, 1.0, 0, 0, 0, 0, 0, 0]])
    o = O.slice_chain(spec, np.array([0.5]), 1.0, 1.0, 1, np.ones((1, 1)), tu, logL_replay=rep)
    assert o["steps"][0] == 2 and o["logL"] == 1.0


# ---------------------------------------------------------------- sample-collection append
def test_replace_index_kats():
    """replace_index (internals/maps.py:15-25) known answers of internals/tests/test_maps.py:7-30: an
    update that would run past the end is shifted back (lax.dynamic_update_slice clamps the start)."""
    op = np.array([0, 1, 2, 3, 4])
    assert list(O.replace_index(op, [5, 5], 0)) == [5, 5, 2, 3, 4]
    assert list(O.replace_index(op, [5], 0)) == [5, 1, 2, 3, 4]
    assert list(O.replace_index(op, [5, 5], 4)) == [0, 1, 2, 5, 5]
    assert list(O.replace_index(op, [5], 4)) == [0, 1, 2, 3, 5]
    assert O.update_start(150, 32, 160) == 128 and O.update_start(100, 32, 160) == 100
