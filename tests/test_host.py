"""Host-side logic (no GPU): C-ABI library, model parsing and parameter packing,
public defaults, termination conditions and the results file format."""
import ctypes as C
import io
import os
import re

import numpy as np
import pytest
from scipy.special import ndtri

import jaxns_amd as J
from jaxns_amd import _lib
from jaxns_amd import distributions as tfpd
from jaxns_amd.engine import RunSpec, flatten_term_cond, round_up_max_samples, round_up_num_live_points

from problems import gmm_problem, linreg_problem, mvn_problem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "hipns.h")).read()
    declared = set(re.findall(r"^\w[\w\s\*]*?\b(hipns_\w+)\(", hdr, flags=re.M))
    assert declared, "no declarations parsed"
    L = _lib.lib()
    for name in sorted(declared):
        assert hasattr(L, name), name
    assert set(_lib.EXPORTED) <= declared
    assert L.hipns_abi_version() == 1


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", "/nonexistent/libhipns.so")
    with pytest.raises(_lib.HipnsError):
        _lib.lib()


def test_struct_layouts():
    assert C.sizeof(_lib.RunDesc) == 88
    assert C.sizeof(_lib.RunState) == 264
    assert C.sizeof(_lib.TermCond) == 10 * 4 + 10 * 8
    assert C.sizeof(_lib.Launch) == 40
    assert C.sizeof(_lib.Buffers) == 168


def test_round_up_kats():
    # nested_samplers/sharded/tests/test_sharded_static.py:4-7
    assert round_up_num_live_points(10, 0.5, 1) == 10
    assert round_up_num_live_points(10, 0.5, 2) == 12
    assert round_up_num_live_points(10, 0.5, 3) == 12
    assert round_up_max_samples(1001, 150, 0) == 1050
    assert round_up_max_samples(1000, 150, 1) == 1200


def test_public_defaults():
    # public.py:69-107 and ShardedStaticNestedSampler.__post_init__ (sharded_static.py:625-647)
    model, _, _, _ = mvn_problem(10)
    ns = J.NestedSampler(model, num_live_points=300, s=5)
    assert ns.num_slices == 50 and ns.k == 0 and ns.c == 300 and ns.num_live_points == 300
    assert ns.max_samples == 300 * 100
    ns2 = J.NestedSampler(model)
    assert ns2.c == 300 and ns2.num_slices == 50
    ns3 = J.NestedSampler(model, difficult_model=True, parameter_estimation=True)
    assert ns3.num_slices == 100 and ns3.k == 10 and ns3.c == 1000
    with pytest.raises(ValueError):
        J.NestedSampler(model, k=50)
    t = ns._term(None)
    assert t["dlogZ"] == pytest.approx(np.log(1 + 1e-3)) and t["max_samples"] == 30000


def test_slice_sampler_codes_and_gradient_slice_scope():
    # UniDimSliceSampler flags (uni_slice_sampler.py:276-320) -> the run descriptor's sampler code
    from jaxns_amd.engine import DeviceBatch, slice_sampler_code
    from jaxns_amd.samplers import UniDimSliceSampler
    assert [slice_sampler_code(gg, gs) for gg, gs in ((0, 0), (1, 0), (0, 1), (1, 1))] == [0, 2, 3, 4]
    model, _, _, _ = mvn_problem(4)
    assert UniDimSliceSampler(model, 8, 1, True, gradient_slice=True)._sampler_code == 3
    assert UniDimSliceSampler(model, 8, 1, True, gradient_slice=True, gradient_guided=True)._sampler_code == 4
    with pytest.raises(NotImplementedError):
        UniDimSliceSampler(model, 8, 1, True, perfect=False)
    # gradient_slice runs in the fused kernel only (the MVN and mixture families)
    lin, _ = linreg_problem(64, 3)[:2]
    with pytest.raises(NotImplementedError):
        UniDimSliceSampler(lin, 8, 0, True, gradient_slice=True)
    with pytest.raises(NotImplementedError):
        DeviceBatch([RunSpec(model=lin, num_live_points=16, num_slices=4, max_samples=1000)], gradient_slice=True)
    cb = J.Model(model.prior_model, lambda x: -0.5 * (x * x).sum(-1))
    with pytest.raises(NotImplementedError):
        DeviceBatch([RunSpec(model=cb, num_live_points=16, num_slices=4, max_samples=1000)], gradient_slice=True)


def test_term_cond_combinators():
    a = J.TerminationCondition(dlogZ=0.1)
    b = J.TerminationCondition(max_samples=10)
    assert isinstance(a & b, J.TerminationConditionConjunction)
    assert isinstance(a | b, J.TerminationConditionDisjunction)
    assert flatten_term_cond(a) == {"dlogZ": 0.1}


def test_model_parsing_and_transform():
    def prior_model():
        x = yield J.Prior(tfpd.Uniform(low=0.0, high=2.0), name="x")
        y = yield J.Prior(tfpd.Normal(loc=[1.0, 2.0], scale=[3.0, 4.0]), name="y")
        return x, y

    m = J.Model(prior_model, J.MultivariateNormalLikelihood(np.zeros(3), cov=np.eye(3)))
    assert m.U_ndims == 3
    U = np.array([[0.25, 0.5, 0.975]])
    X = m.transform(U)
    np.testing.assert_allclose(X["x"], [[0.5]])
    np.testing.assert_allclose(X["y"], [[1.0, 2.0 + 4.0 * ndtri(0.975)]])
    with pytest.raises(J.InvalidPriorName):
        def bad():
            yield J.Prior(tfpd.Normal(0.0, 1.0), name="a")
            yield J.Prior(tfpd.Normal(0.0, 1.0), name="a")
            return None
        J.Model(bad, J.MultivariateNormalLikelihood(np.zeros(2), cov=np.eye(2)))


def test_mvn_likelihood_defers_its_factor_to_the_device():
    """A multivariate-normal likelihood given a covariance keeps it: the Cholesky factor is formed on the
    GPU in FP64 when first needed, so building the model needs no device; a given scale_tril is kept."""
    lik = J.MultivariateNormalLikelihood(np.zeros(3), cov=np.eye(3) * 2.0)
    assert lik._scale_tril is None and lik.cov.shape == (3, 3)
    L = np.array([[1.0, 0, 0], [0.5, 2.0, 0], [0.1, 0.2, 3.0]])
    lik2 = J.MultivariateNormalLikelihood(np.zeros(3), scale_tril=L)
    np.testing.assert_array_equal(lik2.scale_tril, L)
    assert lik2.log_const == pytest.approx(-1.5 * np.log(2 * np.pi) - np.log(6.0))


def test_gmm_params_and_analytic_evidence():
    model, log_Z, means, sig = gmm_problem(2, seed=1)
    # brute-force quadrature of the analytic mixture evidence in D=2
    g = np.linspace(-10, 10, 1201)
    X = np.stack(np.meshgrid(g, g, indexing="ij"), -1).reshape(-1, 2)
    L = model.log_likelihood.log_prob_numpy(X)
    dx = g[1] - g[0]
    bf = np.log(np.sum(np.exp(L)) * dx * dx) - 2 * np.log(20.0)
    assert abs(bf - log_Z) < 1e-3
    prm, _ = model.log_likelihood.device_params(model, 32)
    assert prm.size == (3 + 4) * 32 + 8


def test_linreg_device_packing_reproduces_likelihood():
    """The packed regression block (column-scaled BF16 hi/lo X, FP64-formed r0 = y - X w0, per-run
    p, s, w0) evaluated with the device's BF16x3 arithmetic reproduces log N(y; X w, sigma^2 I)
    (scipy): 1e-5 relative for prior draws (3e-5 with 10% of the columns scaled by 100, where the
    16-bit split of the few dominant weights is shared by every row), and within 1e-3 nat + 1e-6
    relative at posterior draws, where r0 carries the residual."""
    from scipy.stats import norm
    from jaxns_amd.likelihoods import _bf16_rne, _bf16_to_f32
    for scaled in (0.0, 0.1):
        model, _, lik = linreg_problem(300, 20, seed=4, sigma=0.3, tau=1.5, scaled_cols=scaled)
        D, dps = 20, 32
        prm, c = lik.device_params(model, dps)
        blk, npad, panels = lik.data.packed(dps)
        assert npad == 320 and panels == 2 and blk.size == 2 * npad + 2 * npad * 16 * panels
        ne = npad * 16 * panels  # bf16 elements of one X part
        Xf = blk[2 * npad + ne:].reshape(panels, npad, 16).transpose(1, 0, 2).reshape(npad, 16 * panels)
        np.testing.assert_allclose(Xf[:300, :D], lik.X / lik.data.col_scale, rtol=1e-7)
        np.testing.assert_array_equal(blk[npad + ne:npad + ne + 300], lik.y.astype(np.float32))
        r0 = blk[:npad].astype(np.float64)
        from_panels = lambda a: a.reshape(panels, npad, 16).transpose(1, 0, 2).reshape(npad, 16 * panels)
        hi = from_panels(blk[npad:npad + ne // 2].view(np.uint16))
        lo = from_panels(blk[npad + ne // 2:npad + ne].view(np.uint16))
        Xh, Xl = _bf16_to_f32(hi).astype(np.float64), _bf16_to_f32(lo).astype(np.float64)
        p, s_, w0 = prm[dps:2 * dps], prm[2 * dps:3 * dps], prm[3 * dps:4 * dps]
        inv2s2 = float(prm[4 * dps:4 * dps + 2].view(np.float64)[0])
        assert np.all(prm[:D] == 1) and inv2s2 == 0.5 / 0.3 ** 2
        rng = np.random.default_rng(1)
        # prior draws and posterior draws
        A = np.eye(D) / 1.5 ** 2 + lik.X.T @ lik.X / 0.3 ** 2
        m = np.linalg.solve(A, lik.X.T @ lik.y / 0.3 ** 2)
        W = np.concatenate([rng.normal(size=(20, D)) * 1.5, rng.multivariate_normal(m, np.linalg.inv(A), size=20)])
        U = norm.cdf(W / 1.5)
        for i, (u, w) in enumerate(zip(U, W)):
            q = np.zeros(dps, np.float32)
            q[:D] = ndtri(u)
            dw = (np.float32(s_) * q + np.float32(p)).astype(np.float32) - w0
            dh = _bf16_to_f32(_bf16_rne(dw)).astype(np.float64)
            dl = _bf16_to_f32(_bf16_rne((dw - dh).astype(np.float32))).astype(np.float64)
            r = r0 - (Xh @ dh + Xh @ dl + Xl @ dh)
            got = c - np.sum(r * r) * inv2s2
            want = norm.logpdf(lik.y, lik.X @ w, 0.3).sum()
            tol = (3e-5 if scaled else 1e-5) * abs(want) if i < 20 else 1e-6 * abs(want) + 1e-3
            assert abs(got - want) <= tol, (scaled, i, got, want)


def test_linreg_analytic_evidence_and_gradient():
    """log N(y; 0, sigma^2 I + tau^2 X X^T) (Woodbury form) vs scipy's dense logpdf, and its
    closed-form gradient in (log sigma, log tau) vs central finite differences."""
    from scipy.stats import multivariate_normal
    model, log_Z, lik = linreg_problem(60, 5, seed=2, sigma=0.4, tau=0.7)
    dense = multivariate_normal(np.zeros(60), 0.16 * np.eye(60) + 0.49 * lik.X @ lik.X.T).logpdf(lik.y)
    assert abs(log_Z - dense) < 1e-8 * abs(dense)
    g = lik.grad_log_evidence_normal_prior(0.7)
    h = 1e-5
    f = lambda ls, lt: lik.with_sigma(np.exp(ls)).log_evidence_normal_prior(np.exp(lt))
    fd = [(f(np.log(0.4) + h, np.log(0.7)) - f(np.log(0.4) - h, np.log(0.7))) / (2 * h),
          (f(np.log(0.4), np.log(0.7) + h) - f(np.log(0.4), np.log(0.7) - h)) / (2 * h)]
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-6)


def test_runspec_rounding():
    model, _, _, _ = mvn_problem(4)
    s = RunSpec(model=model, num_live_points=301, num_slices=20, k_phantom=2, max_samples=1000)
    assert s.shell == 150 and s.max_samples % (150 * 3) == 0


def _small_results():
    return J.NestedSamplerResults(
        log_Z_mean=-1.5, log_Z_uncert=0.1, ESS=100.0, H_mean=-2.0, samples={"x": np.arange(6.0).reshape(3, 2)},
        parametrised_samples={}, U_samples=np.random.default_rng(0).uniform(size=(3, 2)),
        log_L_samples=np.array([-np.inf, 0.0, 1.0]), log_dp_mean=np.array([-np.inf, -1.0, -0.5]),
        log_X_mean=np.array([-0.1, -0.2, -0.3]), log_posterior_density=np.array([-np.inf, 0.5, 0.25]),
        num_live_points_per_sample=np.array([3, 2, 1]), num_likelihood_evaluations_per_sample=np.array([1, 5, 7]),
        total_num_samples=3, total_phantom_samples=0, total_num_likelihood_evaluations=13, log_efficiency=-1.0,
        termination_reason=4)


def test_results_json_roundtrip(tmp_path):
    """save_results / load_results (utils.py:602-641) write jaxns' namedtuple schema
    (namedtuple_utils.py:34-103): the class path of jaxns' NestedSamplerResults and base64 arrays with
    dtype and shape; every field comes back equal, and its dtype and shape survive."""
    import json
    res = _small_results()
    f = str(tmp_path / "r.json")
    J.save_results(res, f)
    doc = json.load(open(f))
    assert doc["type"] == "__namedtuple__"
    assert doc["__class__"] == "jaxns.nested_samplers.common.types.NestedSamplerResults"
    assert set(doc["__data__"]) == set(res._fields)
    a = doc["__data__"]["U_samples"]
    assert a["type"] == "__ndarray__" and a["__dtype__"] == "float64" and a["__shape__"] == [3, 2]
    back = J.load_results(f)
    assert isinstance(back, J.NestedSamplerResults)
    for k in res._fields:
        a, b = getattr(res, k), getattr(back, k)
        if isinstance(a, dict):
            for kk in a:
                np.testing.assert_array_equal(a[kk], b[kk])
        else:
            np.testing.assert_array_equal(a, b)
            assert np.asarray(a).dtype == np.asarray(b).dtype or np.ndim(a) == 0
    out = J.summary(back)
    assert "logZ=-1.5 +- 0.1" in out and "max(logL)=1.0" in out


def test_loads_a_results_file_in_the_jaxns_schema():
    """A results file laid out as jaxns writes it (jax arrays as __jax_ndarray__, 0-d scalars; fixture
    tests/golden/jaxns_results_schema.json) loads into this package's NestedSamplerResults with the
    fixture's values; written back, it keeps the class path, the fields and every array."""
    import json
    import tempfile
    path = os.path.join(ROOT, "tests", "golden", "jaxns_results_schema.json")
    res = J.load_results(path)
    assert isinstance(res, J.NestedSamplerResults)
    import importlib.util
    spec = importlib.util.spec_from_file_location("mkfix", os.path.join(ROOT, "tests", "golden",
                                                                         "make_results_fixture.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    want = mk.values()
    for k, v in want.items():
        got = getattr(res, k)
        if isinstance(v, dict):
            assert set(got) == set(v)
            for kk in v:
                np.testing.assert_array_equal(got[kk], v[kk])
        else:
            np.testing.assert_array_equal(got, v)
            assert np.shape(got) == np.shape(v)
    assert int(res.total_num_samples) == 6 and float(res.log_Z_mean) == -3.25
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "again.json")
        J.save_results(res, f)
        a, b = json.load(open(path)), json.load(open(f))
        assert b["__class__"] == a["__class__"] and set(b["__data__"]) == set(a["__data__"])
        for k, v in a["__data__"].items():
            if "type" in v:
                assert {kk: v[kk] for kk in ("__dtype__", "__data__", "__shape__")} == \
                    {kk: b["__data__"][k][kk] for kk in ("__dtype__", "__data__", "__shape__")}, k
    J.summary(res)


def test_unknown_namedtuple_loads_without_import(tmp_path):
    """deserialise_namedtuple never imports the stored class path: an unknown NamedTuple comes back
    as a plain namedtuple with the stored fields (the reference's test_gh171 pytree)."""
    import json
    from jaxns_amd.utils import load_pytree
    f = str(tmp_path / "p.json")
    json.dump({"type": "__namedtuple__", "__class__": "some.module.that.does.not.exist.MockPyTree171",
               "__data__": {"x": {"type": "__jax_ndarray__", "__dtype__": "float32",
                                  "__data__": __import__("base64").b64encode(np.array([1., 2., 3.], np.float32)
                                                                            .tobytes()).decode(),
                                  "__shape__": [3]}}}, open(f, "w"))
    p = load_pytree(f)
    assert type(p).__name__ == "MockPyTree171"
    np.testing.assert_array_equal(p.x, np.array([1., 2., 3.], np.float32))


def test_summary_fields():
    """summary (utils.py:284-430): termination text, counts, log Z and max(log L) rounded to the
    uncertainty, H, ESS, and per dimension mean +- std | percentiles | MAP est. | max(L) est."""
    res = _small_results()._replace(termination_reason=4 | (1 << 10))
    out = J.summary(res, f_obj=io.StringIO())
    assert "Small remaining evidence" in out and "no seed points left" in out
    assert "likelihood evals: 13" in out and "samples: 3" in out
    assert "x[#]: mean +- std.dev. | 10%ile / 50%ile / 90%ile | MAP est. | max(L) est." in out
    # MAP point is row 1 (largest log posterior), max(L) point row 2
    line0 = [ln for ln in out.splitlines() if ln.startswith("x[0]:")][0]
    assert line0.split("|")[-2].strip() == "2.0" and line0.split("|")[-1].strip() == "4.0"


def test_sample_evidence_moments():
    """sample_evidence (utils.py:433-480) draws T_i ~ Beta(n_i, 1) and sums Z = sum_i X_{i-1}(1 - T_i) L_i
    (a right-point rule, unlike the trapezoid of compute_evidence_stats).  Its mean has the closed form
    E[Z] = sum_i L_i / (n_i + 1) prod_{j<i} n_j / (n_j + 1)."""
    rng = np.random.default_rng(3)
    M = 400
    logL = np.sort(rng.normal(size=M)) * 3
    n = np.full(M, 20.0)
    n[-20:] = np.arange(20, 0, -1)
    EX_prev = np.concatenate([[1.0], np.cumprod(n / (n + 1.0))[:-1]])
    EZ = np.sum(np.exp(logL) * EX_prev / (n + 1.0))
    Z = np.exp(J.sample_evidence(5, n, logL, S=20000))
    assert abs(Z.mean() - EZ) < 4 * Z.std() / np.sqrt(Z.size)
