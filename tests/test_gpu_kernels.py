"""GPU parity of the libhipns kernels against the CPU oracle (called through the C ABI)."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as O
from jaxns_amd import _lib
from jaxns_amd._lib import check, lib, ptr

from problems import (gmm_oracle_spec, gmm_problem, linreg_oracle_spec, linreg_problem, mvn_oracle_spec,
                      mvn_problem)

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def _stream():
    return _lib.stream_ptr()


def say(*a):
    print(*a, flush=True)


def test_philox_kat():
    """Random123 known-answer vectors for Philox4x32-10."""
    torch = _torch()
    cases = [([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
             ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
             ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
              [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1])]
    for ctr, key, want in cases:
        c = torch.tensor(np.array(ctr, np.uint32).view(np.int32), device="cuda")
        o = torch.empty(4, dtype=torch.int32, device="cuda")
        check(lib().hipns_philox4x32(ptr(c), 1, key[0], key[1], ptr(o), _stream()))
        got = o.cpu().numpy().view(np.uint32)
        assert [int(x) for x in got] == want


def test_ndtri_f32_accuracy():
    """Device q = ndtri against scipy (FP64): relative 2e-6 in the body, absolute 2e-5 in the far tails."""
    from scipy.special import ndtri
    torch = _torch()
    u = np.concatenate([np.linspace(1e-6, 1 - 1e-6, 100001), np.logspace(-38, -1, 3000),
                        1 - np.logspace(-7.2, -1, 1000)]).astype(np.float32)
    u = u[(u > 0) & (u < 1)]
    ud = torch.from_numpy(u).cuda()
    z = torch.empty_like(ud)
    check(lib().hipns_ndtri_f32(ptr(ud), ud.numel(), ptr(z), _stream()))
    got = z.cpu().numpy().astype(np.float64)
    ref = ndtri(u.astype(np.float64))
    err = np.abs(got - ref) / np.maximum(np.abs(ref), 1.0)
    say("ndtri max scaled err", err.max())
    assert err.max() < 2e-6


def test_slice_quotient_matches_ieee_division():
    """The slice bracket's quotient (hardware reciprocal, one Newton step, one residual correction)
    is bit-identical to IEEE binary32 division: every denominator of six whole binades (all 2^23
    mantissas, random signs) against numerators 1 - u0 and -u0, then 2^24 log-uniform pairs over the
    bracket's range |a| in [2^-60, 2], |n| in [2^-94, 1]."""
    torch = _torch()
    rng = np.random.default_rng(11)

    def run(a, n):
        ad, nd = torch.from_numpy(a).cuda(), torch.from_numpy(n).cuda()
        qd = torch.empty_like(ad)
        check(lib().hipns_slice_quotients(ptr(ad), ptr(nd), ad.numel(), ptr(qd), _stream()))
        got = qd.cpu().numpy()
        want = a / n
        bad = np.flatnonzero(got.view(np.uint32) != want.view(np.uint32))
        if bad.size:
            i = bad[0]
            say("mismatch", bad.size, "a", repr(a[i]), "n", repr(n[i]), "got", repr(got[i]), "want", repr(want[i]))
        return bad.size

    mant = np.arange(1 << 23, dtype=np.uint32)
    for e in (-1, -2, -7, -23, -60, -94):
        n = (((127 + e) << 23) | mant).astype(np.uint32).view(np.float32)
        n = np.where(rng.random(n.size) < 0.5, -n, n).astype(np.float32)
        u0 = np.exp2(rng.uniform(-60, 0, size=n.size)).astype(np.float32)
        u0 = np.where(rng.random(n.size) < 0.5, u0, np.float32(1) - u0).astype(np.float32)
        u0 = np.clip(u0, np.float32(2.0 ** -60), np.float32(1 - 2.0 ** -24))
        a = np.where(rng.random(n.size) < 0.5, np.float32(1) - u0, -u0).astype(np.float32)
        assert run(a, n) == 0, e
    M = 1 << 24
    a = (np.exp2(rng.uniform(-60, 1, size=M)) * rng.choice([-1.0, 1.0], size=M)).astype(np.float32)
    n = (np.exp2(rng.uniform(-94, 0, size=M)) * rng.choice([-1.0, 1.0], size=M)).astype(np.float32)
    assert run(a, n) == 0


def test_slice_reciprocal_is_correctly_rounded():
    """The reciprocal the slice bracket's quotient starts from (hardware reciprocal and one Newton step)
    equals 1 / n rounded to nearest for every binary32 n of magnitude in [2^-94, 1] (all 2^23 mantissas
    of 94 binades and n = 1, positive; four binades and n = -1 negative): with y = RN(1/n), one
    residual correction of q = a y gives RN(a / n) (P. Markstein, IBM J. Res. Dev. 34(1), 1990)."""
    torch = _torch()
    mant = torch.arange(1 << 23, dtype=torch.int64, device="cuda")
    y = torch.empty(1 << 23, dtype=torch.float32, device="cuda")
    bad = 0
    cases = [(e, 1.0) for e in range(-94, 0)] + [(e, -1.0) for e in (-1, -2, -23, -94)]
    for e, sign in cases:
        bits = ((127 + e) << 23) + mant
        n = (bits.to(torch.int32).view(torch.float32) * sign).contiguous()
        check(lib().hipns_slice_quotients(None, ptr(n), n.numel(), ptr(y), _stream()))
        got = y.cpu().numpy()
        nn = n.cpu().numpy()
        want = np.float32(1.0) / nn
        k = np.flatnonzero(got.view(np.uint32) != want.view(np.uint32))
        if k.size:
            say("binade", e, sign, "mismatches", k.size, "n", repr(nn[k[0]]), "got", repr(got[k[0]]),
                "want", repr(want[k[0]]))
        bad += k.size
    one = torch.tensor([1.0, -1.0], dtype=torch.float32, device="cuda")
    y1 = torch.empty(2, dtype=torch.float32, device="cuda")
    check(lib().hipns_slice_quotients(None, ptr(one), 2, ptr(y1), _stream()))
    assert y1.cpu().tolist() == [1.0, -1.0]
    assert bad == 0


def _evidence_inputs(M, seed):
    rng = np.random.default_rng(seed)
    logL = np.sort(rng.normal(size=M).cumsum() * 0.01 + rng.normal(size=M) * 0.1)
    logL[: min(5, M)] = -np.inf if M > 10 else logL[: min(5, M)]
    n = rng.integers(1, 300, size=M).astype(np.float64)
    return logL, n


@pytest.mark.parametrize("M,stop", [(1, None), (7, None), (2048, None), (5000, None), (20000, 12345), (70001, None)])
def test_evidence_scan_vs_oracle(M, stop):
    torch = _torch()
    say(f"evidence_scan M={M} stop={stop}")
    logL, n = _evidence_inputs(M, M)
    fin_o, per_o = O.compute_evidence_stats(logL, n, stop)
    L = torch.from_numpy(logL).cuda()
    N = torch.from_numpy(n).cuda()
    per = torch.empty((8, M), dtype=torch.float64, device="cuda")
    fin = torch.empty(8, dtype=torch.float64, device="cuda")
    ws = torch.empty(lib().hipns_evidence_scan_workspace_size(M), dtype=torch.uint8, device="cuda")
    check(lib().hipns_evidence_scan(ptr(L), ptr(N), M, -1 if stop is None else stop, ptr(per), ptr(fin), ptr(ws),
                                    ws.numel(), _stream()))
    per = per.cpu().numpy().T
    fin = fin.cpu().numpy()
    # same values up to FP64 reassociation of the log-space sums
    for a, b in ((per, per_o), (fin[None], fin_o[None])):
        fa, fb = np.isfinite(a), np.isfinite(b)
        assert np.array_equal(fa, fb)
        assert np.array_equal(a[~fa], b[~fb])
        np.testing.assert_allclose(a[fa], b[fb], rtol=1e-9, atol=1e-9)


def test_evidence_scan_closed_form_large():
    """3 000 000 samples (1465 tiles of 2048, so the tile scan composes two tiles per thread) against
    the closed forms of a constant likelihood L = 1 with n = N throughout: log X_i = i log(N/(N+1)),
    log X2_i = i log(N/(N+2)), and Z_M = (1 - X_1)/2 + X_1 - X_M (the initial carry has log L = -inf,
    so the first mean likelihood is 1/2)."""
    torch = _torch()
    M, N = 3_000_000, 4096.0
    say(f"evidence_scan closed form M={M}")
    L = torch.zeros(M, dtype=torch.float64, device="cuda")
    n = torch.full((M,), N, dtype=torch.float64, device="cuda")
    per = torch.empty((8, M), dtype=torch.float64, device="cuda")
    fin = torch.empty(8, dtype=torch.float64, device="cuda")
    ws = torch.empty(lib().hipns_evidence_scan_workspace_size(M), dtype=torch.uint8, device="cuda")
    check(lib().hipns_evidence_scan(ptr(L), ptr(n), M, -1, ptr(per), ptr(fin), ptr(ws), ws.numel(), _stream()))
    i = np.array([1, 2, 2047, 2048, 2049, 1024 * 2048 - 1, 1024 * 2048, 2_500_001, M], dtype=np.int64)
    logX = per[1].cpu().numpy()[i - 1]
    logX2 = per[2].cpu().numpy()[i - 1]
    np.testing.assert_allclose(logX, i * np.log(N / (N + 1.0)), rtol=1e-9)
    np.testing.assert_allclose(logX2, i * np.log(N / (N + 2.0)), rtol=1e-9)
    X1 = N / (N + 1.0)
    Z = 0.5 * (1.0 - X1) + X1 - np.exp(M * np.log(X1))
    f = fin.cpu().numpy()
    np.testing.assert_allclose(f[3], np.log(Z), rtol=0, atol=1e-10)  # Z is close to 1: absolute in log Z
    np.testing.assert_allclose(f[1], M * np.log(X1), rtol=1e-9)


def test_argsort_stable_vs_numpy():
    torch = _torch()
    rng = np.random.default_rng(3)
    for n in [1, 5, 2048, 2049, 10000, 100003]:
        say(f"argsort n={n}")
        keys = np.round(rng.normal(size=n), 1)  # many ties
        keys[rng.integers(0, n, size=max(1, n // 50))] = -np.inf
        keys[rng.integers(0, n, size=max(1, n // 50))] = np.inf
        k = torch.from_numpy(keys).cuda()
        p = torch.empty(n, dtype=torch.int64, device="cuda")
        ws = torch.empty(lib().hipns_argsort_workspace_size(n), dtype=torch.uint8, device="cuda")
        check(lib().hipns_argsort_stable(ptr(k), n, ptr(p), ptr(ws), ws.numel(), _stream()))
        assert np.array_equal(p.cpu().numpy(), np.argsort(keys, kind="stable"))


def test_argsort_stable_edge_keys():
    """Keys where an order of bit patterns and the comparison order part: -0 and +0 tie (input order
    kept), NaNs of either sign and any payload tie after +inf, subnormals and both signs of every
    magnitude, all-equal keys, and sizes spanning many 2048-element tiles with ragged ends."""
    torch = _torch()
    rng = np.random.default_rng(17)

    def run(keys):
        n = keys.size
        k = torch.from_numpy(keys).cuda()
        p = torch.empty(n, dtype=torch.int64, device="cuda")
        ws = torch.empty(lib().hipns_argsort_workspace_size(n), dtype=torch.uint8, device="cuda")
        check(lib().hipns_argsort_stable(ptr(k), n, ptr(p), ptr(ws), ws.numel(), _stream()))
        got = p.cpu().numpy()
        assert np.array_equal(got, np.argsort(keys, kind="stable")), f"n={n}"

    specials = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, -np.nan, 5e-324, -5e-324, 2.2e-308, -2.2e-308,
                         1.7976931348623157e308, -1.7976931348623157e308, 1.0, -1.0])
    nan_payload = np.array([0x7FF0000000000001, 0xFFF8000000000123, 0x7FFFFFFFFFFFFFFF], dtype=np.uint64).view(np.float64)
    for n in [3, 2047, 4097, 65536 + 11, (1 << 21) + 3]:
        say(f"argsort edge keys n={n}")
        keys = rng.choice(np.concatenate([specials, nan_payload]), size=n)
        mask = rng.random(n) < 0.5
        keys[mask] = rng.normal(size=int(mask.sum())) * 10.0 ** rng.integers(-300, 300, size=int(mask.sum()))
        run(keys)
    say("argsort all-equal keys")
    run(np.full(10000, -0.0))
    run(np.zeros(4096 * 3 + 1))
    say("argsort large ragged")
    keys = np.round(rng.normal(size=(1 << 22) + 1234), 2)
    run(keys)


def _tree(n, seed):
    rng = np.random.default_rng(seed)
    logL = [0.0]
    parent = []
    for _ in range(n):
        i = int(rng.integers(0, len(logL)))
        logL.append(logL[i] + float(rng.uniform(0.0, 1.0)))
        parent.append(i)
    return np.array(parent, np.int64), np.array(logL[1:])


@pytest.mark.parametrize("n,num_samples", [(6, None), (200, None), (50000, None), (3000, 2500)])
def test_count_crossed_edges_vs_oracle(n, num_samples):
    torch = _torch()
    say(f"count_crossed_edges n={n}")
    sender, logL = _tree(n, n)
    if num_samples is not None:
        sender[num_samples:] = 0
        logL[num_samples:] = np.inf
    si_o, nl_o = O.count_crossed_edges(sender, logL, num_samples)
    s = torch.from_numpy(sender.astype(np.int32)).cuda()
    L = torch.from_numpy(logL).cuda()
    si = torch.empty(n, dtype=torch.int64, device="cuda")
    nl = torch.empty(n, dtype=torch.int64, device="cuda")
    ws = torch.empty(lib().hipns_count_crossed_edges_workspace_size(n), dtype=torch.uint8, device="cuda")
    check(lib().hipns_count_crossed_edges(ptr(s), ptr(L), n, -1 if num_samples is None else num_samples, ptr(si),
                                          ptr(nl), ptr(ws), ws.numel(), _stream()))
    assert np.array_equal(si.cpu().numpy(), si_o)
    assert np.array_equal(nl.cpu().numpy(), nl_o)


def _batch(model, N=64, num_slices=None, k=0, seed=1, lpc=2, max_ms=None, wpb=2):
    from jaxns_amd.engine import DeviceBatch, RunSpec
    D = model.U_ndims
    spec = RunSpec(model=model, num_live_points=N, num_slices=num_slices or 2 * D, k_phantom=k,
                   max_samples=max_ms or N * 20, seed=seed)
    return DeviceBatch([spec], lanes_per_chain=lpc, waves_per_block=wpb, max_steps=64)


def _linreg_posterior_U(lik, tau, n, rng):
    """U of n draws from the analytic posterior of w ~ N(0, tau^2 I)."""
    from scipy.stats import norm
    A = np.eye(lik.D) / tau ** 2 + lik.X.T @ lik.X / lik.sigma ** 2
    m = np.linalg.solve(A, lik.X.T @ lik.y / lik.sigma ** 2)
    W = rng.multivariate_normal(m, np.linalg.inv(A), size=n)
    return norm.cdf(W / tau)


@pytest.mark.parametrize("fam,D,lpc,wpb,nd", [("mvn", 10, 2, 2, 0), ("mvn", 32, 2, 2, 0), ("mvn", 32, 4, 2, 0),
                                              ("mvn", 70, 2, 2, 0), ("gmm", 7, 2, 2, 0), ("gmm", 40, 4, 2, 0),
                                              ("gmm", 7, 8, 2, 0), ("gmm", 40, 8, 1, 0), ("gmm", 100, 8, 2, 0),
                                              ("linreg", 16, 2, 4, 4096), ("linreg", 16, 4, 1, 4096),
                                              ("linreg", 5, 2, 3, 37), ("linreg", 100, 4, 2, 1000)])
def test_batched_loglik_vs_oracle(fam, D, lpc, wpb, nd):
    """Model.forward on the device (FP32; BF16x3 MFMA for the regression rows) vs the FP64
    oracle: 1e-5 relative."""
    torch = _torch()
    say(f"loglik {fam} D={D} lpc={lpc} wpb={wpb} ndata={nd}")
    rng = np.random.default_rng(0)
    B = 1000
    U = rng.uniform(0.02, 0.98, size=(B, D)).astype(np.float32)
    if fam == "mvn":
        model, _, mu, Sig = mvn_problem(D, seed=D)
        spec = mvn_oracle_spec(D, mu, Sig)
    elif fam == "gmm":
        model, _, means, sig = gmm_problem(D, seed=D)
        spec = gmm_oracle_spec(D, means, sig)
        # points near the modes so that the mixture is not -inf in FP32
        U = ((means[rng.integers(0, means.shape[0], B)] + rng.normal(size=(B, D)) * 0.5 + 10) / 20).clip(0.01, 0.99)
        U = U.astype(np.float32)
    else:
        model, _, lik = linreg_problem(nd, D, seed=D, sigma=0.3, tau=1.0)
        spec = linreg_oracle_spec(lik, 1.0)
        U[B // 2:] = _linreg_posterior_U(lik, 1.0, B - B // 2, rng).clip(1e-6, 1 - 1e-6)
    b = _batch(model, lpc=lpc, wpb=wpb)
    Up = np.full((B, b.dps), 0.5, np.float32)
    Up[:, :D] = U
    Ud = torch.from_numpy(Up).cuda()
    ref = spec.log_likelihood(U.astype(np.float64))
    for terms in ((1, 3, 4) if fam == "linreg" else (0,)):
        out = torch.empty(B, dtype=torch.float64, device="cuda")
        b.launch.lr_terms = terms
        check(lib().hipns_log_likelihood(ptr(Ud), B, ptr(out), ptr(b.runs_dev), b.runs_host, C.byref(b.bufs),
                                         C.byref(b.launch), _stream()))
        got = out.cpu().numpy()
        rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1.0)
        say("terms", terms, "max rel err", rel.max())
        # BF16x3 on prior draws: the 16-bit split of w - w0 is shared by all rows, 3e-5 relative;
        # every other form (plain FP32 = 1, four products = 4, MVN, GMM): 1e-5 relative
        assert rel.max() < (3e-5 if terms == 3 else 1e-5)
        if fam == "linreg":  # posterior draws: the centred residual keeps the error far below the posterior scale
            post = np.abs(got - ref)[B // 2:]
            say("posterior draws: max abs err", post.max())
            assert post.max() < 1e-6 * np.abs(ref[B // 2:]).max() + 2e-3


@pytest.mark.parametrize("fam,D,lpc,k,wpb,faces", [("mvn", 10, 2, 0, 2, False), ("mvn", 32, 2, 2, 2, False),
                                                   ("mvn", 32, 4, 0, 2, False), ("mvn", 5, 4, 1, 2, False),
                                                   ("linreg", 16, 2, 1, 4, False), ("linreg", 40, 4, 0, 3, False),
                                                   ("mvn", 10, 2, 0, 2, True), ("mvn", 5, 4, 0, 1, True),
                                                   ("gmm", 40, 8, 1, 2, False), ("gmm", 100, 8, 0, 1, False),
                                                   ("gmm", 6, 4, 0, 2, True), ("mvn", 70, 4, 0, 2, False)])
def test_slice_chains_injected_bit_exact(fam, D, lpc, k, wpb, faces):
    """UniDimSliceSampler chains with injected draws: the device trace (t of every pick),
    the accept/reject sequence, eval counts and the accepted points are bit-exact with
    the binary32 oracle replaying the device's likelihood values; the device likelihood
    itself agrees with the FP64 oracle to 1e-5.  faces=True seeds chains at points with a
    coordinate within 2^-60 of a face of the unit cube (down to the smallest subnormal), where the
    bracket takes the general candidate rule, and just inside that range."""
    torch = _torch()
    say(f"slice injected {fam} D={D} lpc={lpc} k={k} wpb={wpb} faces={faces}")
    if fam == "mvn":
        model, _, mu, Sig = mvn_problem(D, seed=100 + D)
        spec = mvn_oracle_spec(D, mu, Sig)
    elif fam == "gmm":
        model, _, means, sig = gmm_problem(D, seed=100 + D)
        spec = gmm_oracle_spec(D, means, sig)
    else:
        model, _, lik = linreg_problem(300, D, seed=100 + D, sigma=0.5, tau=1.0)
        spec = linreg_oracle_spec(lik, 1.0)
    N = 64
    ns = 2 * D
    b = _batch(model, N=N, num_slices=ns, k=k, seed=7, lpc=lpc, wpb=wpb)
    b.init_state()
    torch.cuda.synchronize()
    # (a coordinate exactly on a face collapses the bracket to t = 0, which the reference's loop never
    # leaves unless the seed itself is accepted, so the cases stop just inside the faces)
    face_vals = [1e-30, 2.0 ** -70, 2.0 ** -59, 2.0 ** -149, 1.0 - 2.0 ** -24, 2.0 ** -61]
    if faces:
        for i, v in enumerate(face_vals):
            b.live_U[0][N - 1 - i, i % D] = v
        torch.cuda.synchronize()
    live_U = b.live_U[0][:N, :D].cpu().numpy()
    live_L = b.live_logL[0][:N].cpu().numpy()
    # near-face seeds no longer satisfy the contour they were drawn under; with contour = -inf every
    # slice accepts its first pick, so each slice tests one bracket (the first from the face point)
    # (the mixture's region above a contour is a few small balls: a lower contour keeps its slices
    # within the step budget)
    contour = -np.inf if faces else float(live_L[(N // 4 if fam == "gmm" else N // 2) - 1])
    rng = np.random.default_rng(D + lpc)
    Cn = N // 2
    seeds = np.array([O.seed_index(live_L, contour, u) for u in rng.uniform(size=Cn)], np.int32)
    if faces:
        seeds[:len(face_vals)] = N - 1 - np.arange(len(face_vals))
    max_steps = 256 if fam == "gmm" else 64
    z = rng.normal(size=(Cn, ns, D))
    dirs = (z / np.linalg.norm(z, axis=-1, keepdims=True)).astype(np.float32)
    dp = np.zeros((Cn, ns, b.dps), np.float32)
    dp[..., :D] = dirs
    tu = rng.uniform(size=(Cn, ns, max_steps)).astype(np.float32)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    sd, dd, td = dev(seeds), dev(dp), dev(tu)
    tr_t = torch.full((Cn, ns, max_steps), float("nan"), dtype=torch.float32, device="cuda")
    tr_L = torch.full((Cn, ns, max_steps), float("nan"), dtype=torch.float64, device="cuda")
    steps = torch.zeros((Cn, ns), dtype=torch.int32, device="cuda")
    kk = max(k, 1)
    phU = torch.zeros((Cn, kk, b.dps), dtype=torch.float32, device="cuda")
    phL = torch.zeros((Cn, kk), dtype=torch.float64, device="cuda")
    launch = _lib.Launch.from_buffer_copy(b.launch)
    launch.max_steps = max_steps
    check(lib().hipns_slice_chains_injected(ptr(b.runs_dev), b.runs_host, C.byref(b.bufs), C.byref(launch), 0, Cn,
                                            contour, ptr(sd), ptr(dd), ptr(td), ptr(tr_t), ptr(tr_L), ptr(steps),
                                            ptr(phU), ptr(phL), _stream()))
    torch.cuda.synchronize()
    g_U = b.new_U[:Cn, :D].cpu().numpy()
    g_L = b.new_logL[:Cn].cpu().numpy()
    g_E = b.new_evals[:Cn].cpu().numpy()
    tr_t, tr_L, steps = tr_t.cpu().numpy(), tr_L.cpu().numpy(), steps.cpu().numpy()
    phU, phL = phU.cpu().numpy(), phL.cpu().numpy()
    for c in range(Cn):
        o = O.slice_chain(spec, live_U[seeds[c]], live_L[seeds[c]], contour, ns, dirs[c], tu[c], k_phantom=k,
                          dtype=np.float32, logL_replay=tr_L[c])
        assert np.array_equal(o["steps"], steps[c]), c
        m = ~np.isnan(o["trace_t"])
        assert np.array_equal(o["trace_t"][m], tr_t[c][m]), c
        assert np.array_equal(o["U"], g_U[c]), c
        assert o["logL"] == g_L[c]
        assert o["evals"] == g_E[c]
        if k:
            assert np.array_equal(o["phantom_U"], phU[c, :k, :D])
            assert np.array_equal(o["phantom_logL"], phL[c, :k])
        # the device likelihood values vs the FP64 oracle at the same (float32) points
        o64 = O.slice_chain(spec, live_U[seeds[c]], live_L[seeds[c]], contour, ns, dirs[c], tu[c], k_phantom=k,
                            dtype=np.float32, logL_replay=None)
        if np.array_equal(o64["steps"], o["steps"]):
            mm = ~np.isnan(o64["trace_logL"])
            ref, got = o64["trace_logL"][mm], tr_L[c][mm]
            fin = np.isfinite(ref)
            assert np.array_equal(ref[~fin], got[~fin])  # points on a face: logL = -inf on both sides
            rel = np.abs(ref[fin] - got[fin]) / np.maximum(np.abs(ref[fin]), 1.0)
            assert rel.size == 0 or rel.max() < 1e-5
    # accepted points satisfy the contour (uni_slice_sampler.py:160-166)
    assert np.all(g_L > contour)
    if faces:
        assert np.all(steps == 1)


def test_full_run_mvn_d10_n300_logZ():
    """Plain single run, D=10 MVN, N=300, s=5: log Z vs analytic within 4 sigma + 0.1."""
    import jaxns_amd as J
    say("full run, MVN D=10, N=300")
    model, log_Z, _, _ = mvn_problem(10, seed=0)
    ns = J.NestedSampler(model, num_live_points=300, s=5, k=0, max_samples=300 * 200)
    reason, state = ns(1)
    res = ns.to_results(reason, state)
    say("logZ", res.log_Z_mean, "+-", res.log_Z_uncert, "true", log_Z, "reason", reason, "evals",
        res.total_num_likelihood_evaluations)
    assert np.isfinite(res.log_Z_mean)
    assert abs(res.log_Z_mean - log_Z) < 4 * res.log_Z_uncert + 0.1
    assert reason & 4  # dlogZ


def test_batched_runs_match_single_runs():
    """A run inside a batch (graph-captured shells) is bit-identical to the same seed run alone."""
    import jaxns_amd as J
    say("batch vs single")
    model, log_Z, _, _ = mvn_problem(8, seed=3)
    ns = J.NestedSampler(model, num_live_points=128, s=3, max_samples=128 * 200)
    r1, s1 = ns(11)
    a = ns.to_results(r1, s1)
    rs, sts = ns.run_batch([10, 11, 12], graph_iters=4)
    b = ns.to_results(rs[1], sts[1])
    assert rs[1] == r1
    assert np.array_equal(a.log_L_samples, b.log_L_samples)
    assert np.array_equal(a.U_samples, b.U_samples)
    assert a.log_Z_mean == b.log_Z_mean


def test_mixture_runs_do_not_depend_on_lane_layout():
    """A mixture run's samples and log Z do not depend on the lanes per chain. The draws are keyed by
    chain, shell, slice and dimension, min / max reductions do not depend on order, and the
    squared-distance partial sums meet in the same pairing tree at every layout. That holds for runs
    of up to 16 dimensions at 2, 4 and 8 lanes per chain and up to 32 at 4 and 8. A run of D = 40 in
    the batch gives the U rows a stride of 64, which 8 lanes per chain need."""
    from jaxns_amd.engine import DeviceBatch, RunSpec
    say("mixture runs across lane layouts")
    Ds = [5, 12, 16, 27, 40]
    models = [gmm_problem(D, seed=70 + D)[0] for D in Ds]
    term = dict(dlogZ=math.log(1.0 + 1e-2), max_samples=1e18)
    out = {}
    for lpc in (2, 4, 8):
        specs = [RunSpec(model=m, num_live_points=96, num_slices=3 * D, max_samples=96 * 150, seed=40 + i)
                 for i, (m, D) in enumerate(zip(models, Ds))]
        b = DeviceBatch(specs, lanes_per_chain=lpc, waves_per_block=2)
        b.run(term, graph_iters=4)
        assert int(b.cap_hits().sum()) == 0
        out[lpc] = [b.to_results(r) for r in range(b.R)]
    for lpcs, runs in (((2, 4, 8), (0, 1, 2)), ((4, 8), (3,))):
        for r in runs:
            a = out[lpcs[0]][r]
            for lpc in lpcs[1:]:
                c = out[lpc][r]
                assert a.total_num_samples == c.total_num_samples, (Ds[r], lpc)
                assert np.array_equal(a.log_L_samples, c.log_L_samples), (Ds[r], lpc)
                assert np.array_equal(a.U_samples, c.U_samples), (Ds[r], lpc)
                assert a.log_Z_mean == c.log_Z_mean, (Ds[r], lpc)


def test_runs_do_not_depend_on_waves_per_block():
    """The autotuner picks the wavefronts per workgroup for each batch; that choice only moves chains
    between workgroups (a run's chains fill whole workgroups, the slots beyond num_live_points / 2
    stay empty), so every run's samples, evaluation counts and log Z are bit-identical at 1, 2 and 4
    wavefronts per workgroup. MVN runs with 150 chains (5, 3 and 2 workgroups of 32, 64 and 128 chain
    slots at 2 lanes per chain) and a mixture batch at 8 lanes per chain."""
    from jaxns_amd.engine import DeviceBatch, RunSpec
    say("runs across waves per workgroup")
    mvn, _, _, _ = mvn_problem(10, seed=21)
    mix = [gmm_problem(D, seed=80 + D)[0] for D in (6, 23)]
    term = dict(dlogZ=math.log(1.0 + 1e-2), max_samples=1e18)
    cases = [("mvn", [RunSpec(model=mvn, num_live_points=300, num_slices=20, max_samples=300 * 120, seed=60 + i)
                      for i in range(3)], 2),
             ("gmm", [RunSpec(model=m, num_live_points=100 + 60 * i, num_slices=2 * m.U_ndims,
                              max_samples=220 * 150, seed=70 + i) for i, m in enumerate(mix)], 8)]
    for name, specs, lpc in cases:
        out = {}
        for wpb in (1, 2, 4):
            b = DeviceBatch(specs, lanes_per_chain=lpc, waves_per_block=wpb)
            b.run(term, graph_iters=4)
            assert int(b.cap_hits().sum()) == 0
            out[wpb] = [b.to_results(r) for r in range(b.R)]
        for r in range(len(specs)):
            a = out[1][r]
            for wpb in (2, 4):
                c = out[wpb][r]
                assert a.total_num_samples == c.total_num_samples, (name, r, wpb)
                assert a.total_num_likelihood_evaluations == c.total_num_likelihood_evaluations, (name, r, wpb)
                assert np.array_equal(a.log_L_samples, c.log_L_samples), (name, r, wpb)
                assert np.array_equal(a.U_samples, c.U_samples), (name, r, wpb)
                assert a.log_Z_mean == c.log_Z_mean, (name, r, wpb)


def test_concurrent_batches_match_sequential():
    """engine.run_concurrent (two batches on streams of their own, their shell graphs replayed in turn)
    gives every run the same samples, evaluation counts and log Z as DeviceBatch.run of each batch
    alone."""
    from jaxns_amd.engine import DeviceBatch, RunSpec, run_concurrent
    say("concurrent batches vs sequential")
    model, _, _, _ = mvn_problem(8, seed=3)

    def make(seeds):
        specs = [RunSpec(model=model, num_live_points=128, num_slices=16, k_phantom=0, max_samples=128 * 100,
                         seed=s) for s in seeds]
        return DeviceBatch(specs, lanes_per_chain=4, waves_per_block=2)

    term = dict(dlogZ=math.log(1.0 + 1e-3), max_samples=1e18)
    con = [make([1, 2, 3]), make([4, 5])]
    run_concurrent(con, term, graph_iters=4)
    seq = [make([1, 2, 3]), make([4, 5])]
    for b in seq:
        b.run(term, graph_iters=4)
    for a, b in zip(con, seq):
        assert np.array_equal(a.log_Z(), b.log_Z())
        for r in range(a.R):
            ra, rb = a.to_results(r), b.to_results(r)
            assert ra.total_num_samples == rb.total_num_samples
            assert np.array_equal(ra.log_L_samples, rb.log_L_samples)
            assert np.array_equal(ra.U_samples, rb.U_samples)
            assert ra.total_num_likelihood_evaluations == rb.total_num_likelihood_evaluations


def test_concurrent_batches_of_different_families():
    """run_concurrent with a multivariate-normal batch (D=8) and a mixture batch (D=5, uniform prior)
    on their two streams, each with its own termination condition: both bit-identical to lone runs."""
    from jaxns_amd.engine import DeviceBatch, RunSpec, run_concurrent
    say("concurrent batches of two likelihood families")
    mvn, _, _, _ = mvn_problem(8, seed=6)
    gmm, _, _, _ = gmm_problem(5, seed=7)

    def make():
        a = DeviceBatch([RunSpec(model=mvn, num_live_points=96, num_slices=16, k_phantom=1, max_samples=96 * 120,
                                 seed=s) for s in (11, 12)], lanes_per_chain=4, waves_per_block=2)
        b = DeviceBatch([RunSpec(model=gmm, num_live_points=160, num_slices=10, k_phantom=0,
                                 max_samples=160 * 150, seed=s) for s in (13, 14, 15)], lanes_per_chain=2,
                        waves_per_block=2)
        return [a, b]

    terms = [dict(dlogZ=math.log(1.0 + 1e-3), max_samples=1e18), dict(dlogZ=math.log(1.0 + 1e-2), max_samples=1e18)]
    con = make()
    run_concurrent(con, terms, graph_iters=4)
    seq = make()
    for b, t in zip(seq, terms):
        b.run(t, graph_iters=4)
    for a, b in zip(con, seq):
        assert np.array_equal(a.log_Z(), b.log_Z())
        for r in range(a.R):
            ra, rb = a.to_results(r), b.to_results(r)
            assert ra.total_num_samples == rb.total_num_samples
            assert np.array_equal(ra.log_L_samples, rb.log_L_samples)
            assert ra.total_num_likelihood_evaluations == rb.total_num_likelihood_evaluations


def test_linreg_fp32_refinement():
    """FP32 refinement of accepted points (regression family in BF16x3): after a shell, refine_new
    overwrites every run's accepted-point logL with the plain FP32 form, which agrees with the FP64
    likelihood at the same points to 1e-6 relative (the BF16x3 values themselves to 3e-5); a full run
    with refinement gives log Z within the stated uncertainty of the analytic evidence."""
    from scipy.special import ndtri
    from jaxns_amd._lib import make_term_cond
    from jaxns_amd.engine import DeviceBatch, RunSpec
    torch = _torch()
    say("linreg fp32 refinement")
    model, log_Z, lik = linreg_problem(2048, 16, seed=5, sigma=0.3, tau=1.0)
    specs = [RunSpec(model=model, num_live_points=256, num_slices=5 * 16, k_phantom=0, max_samples=256 * 300,
                     seed=sd) for sd in (1, 2)]
    b = DeviceBatch(specs, lanes_per_chain=2, waves_per_block=4, lr_terms=3, refine_fp32=True)
    b.init_state()
    s = _stream()
    args = (ptr(b.runs_dev), ptr(b.states_dev), b.R, b.runs_host, C.byref(b.bufs))
    tc = make_term_cond(max_samples=1e18)
    for shell in range(3):
        check(lib().hipns_shell_discard(*args, s))
        check(lib().hipns_shell_sample(*args, C.byref(b.launch), s))
        torch.cuda.synchronize()
        raw = b.new_logL.clone()
        b.refine_new()
        torch.cuda.synchronize()
        for r in range(b.R):
            d = b.runs_host[r]
            off, n = int(d.chain_off), int(d.shell)
            U = b.new_U[off:off + n, :16].double().cpu().numpy()
            ref = lik.log_prob_numpy(1.0 * ndtri(U))
            got = b.new_logL[off:off + n].cpu().numpy()
            bf = raw[off:off + n].cpu().numpy()
            rel = np.abs(got - ref) / np.abs(ref)
            say("shell", shell, "run", r, "refined max rel", rel.max(), "bf16x3 max rel", (np.abs(bf - ref) / np.abs(ref)).max())
            assert rel.max() < 1e-6
            assert (np.abs(bf - ref) / np.abs(ref)).max() < 3e-5
        check(lib().hipns_shell_replace(*args, s))
        check(lib().hipns_shell_register(*args, C.byref(tc), s))
    torch.cuda.synchronize()
    full = DeviceBatch(specs[:1], lanes_per_chain=2, waves_per_block=4, lr_terms=3, refine_fp32=True)
    full.run(dict(dlogZ=math.log(1.0 + 1e-3), max_samples=1e18), graph_iters=4)
    z = full.log_Z()
    say("refined run log Z", z[0], "analytic", log_Z)
    assert abs(z[0, 0] - log_Z) < 4 * z[0, 1] + 0.2


@pytest.mark.parametrize("lpc", [4, 8])
def test_uneven_batched_gmm_runs_logZ(lpc):
    """Runs of uneven dimension and live-point count (mixture likelihood, uniform box prior)
    advanced in one batch: each run's log Z against its own analytic value (4 and 8 lanes per
    chain)."""
    from jaxns_amd.engine import DeviceBatch, RunSpec
    say(f"uneven GMM batch lpc={lpc}")
    rng = np.random.default_rng(123)
    specs, truths = [], []
    for r in range(8):
        # 8 lanes per chain run where the row stride (64 here) covers the 8-slot instance's 64 dims; at
        # those dimensions one component (four would lose modes, a bias of the algorithm itself)
        D = int(rng.integers(2, 9)) if lpc == 4 else int(rng.integers(33, 61))
        N = int(rng.integers(128, 320))
        model, log_Z, _, _ = gmm_problem(D, seed=1000 + r, K=4 if lpc == 4 else 1)
        specs.append(RunSpec(model=model, num_live_points=N, num_slices=5 * D, max_samples=N * 400, seed=r))
        truths.append(log_Z)
    b = DeviceBatch(specs, lanes_per_chain=lpc, waves_per_block=2)
    b.run(dict(dlogZ=math.log(1.0 + 1e-3), max_samples=1e18), graph_iters=4)
    z = b.log_Z()
    for r in range(8):
        say(r, specs[r].model.U_ndims, specs[r].num_live_points, z[r], truths[r])
        assert abs(z[r, 0] - truths[r]) < 4 * z[r, 1] + 0.25
    sts = b.states()
    assert all(st.termination_reason & 4 for st in sts)


def test_linreg_hyperparameter_batch_logZ():
    """Bayesian linear regression runs at four (sigma, tau) settings sharing one data set in one
    batch (split-row BF16x3 likelihood): each log Z against log N(y; 0, sigma^2 I + tau^2 X X^T)."""
    from jaxns_amd.engine import DeviceBatch, RunSpec
    say("linreg hyperparameter batch")
    _, _, lik0 = linreg_problem(1024, 8, seed=5, sigma=0.3)
    specs, truths = [], []
    for r, (sg, tau) in enumerate([(0.3, 1.0), (0.5, 1.0), (0.3, 2.0), (0.2, 0.5)]):
        model, log_Z, _ = linreg_problem(0, 0, sigma=sg, tau=tau, data=lik0.data)
        specs.append(RunSpec(model=model, num_live_points=256, num_slices=4 * 8, max_samples=256 * 400, seed=r))
        truths.append(log_Z)
    b = DeviceBatch(specs, lanes_per_chain=2, waves_per_block=4)
    b.run(dict(dlogZ=math.log(1.0 + 1e-3), max_samples=1e18), graph_iters=4)
    z = b.log_Z()
    for r in range(4):
        say(r, z[r], truths[r])
        assert abs(z[r, 0] - truths[r]) < 4 * z[r, 1] + 0.2
    assert all(st.termination_reason & 4 for st in b.states())


@pytest.mark.parametrize("nruns,ms_mult", [(5, 200), (1, 400)])
def test_batch_evidence_matches_per_run(nruns, ms_mult):
    """hipns_batch_evidence (all runs in one set of launches) against the single-run primitives
    (count_crossed_edges + evidence scan per run), on uneven runs with phantom samples (k=2), so
    the samples are not already sorted: phantom counts exact, evidence statistics to 1e-12.  The
    one-run batch has a sort segment of more than 16 tiles of which the run fills only some."""
    import torch
    from jaxns_amd.engine import DeviceBatch, RunSpec
    say(f"batched evidence vs per run: {nruns} runs, max_samples = {ms_mult} N")
    rng = np.random.default_rng(77)
    specs = []
    for r in range(nruns):
        D = int(rng.integers(2, 7))
        N = int(rng.integers(100, 300))
        model, _, _, _ = gmm_problem(D, seed=500 + r)
        specs.append(RunSpec(model=model, num_live_points=N, num_slices=4 * D, k_phantom=2,
                             max_samples=N * ms_mult, seed=r))
    b = DeviceBatch(specs, lanes_per_chain=4, waves_per_block=2)
    b.run(dict(dlogZ=math.log(1.0 + 1e-2), max_samples=1e18), graph_iters=4)
    z1, fin1, ph1 = b.log_Z_per_run()
    fins = torch.empty((b.R, 8), dtype=torch.float64, device="cuda")
    phs = torch.empty(b.R, dtype=torch.int64, device="cuda")
    ws = torch.empty(int(lib().hipns_batch_evidence_workspace_size(b.R, b.runs_host)), dtype=torch.uint8,
                     device="cuda")
    check(lib().hipns_batch_evidence(ptr(b.runs_dev), ptr(b.states_dev), b.R, b.runs_host, C.byref(b.bufs),
                                     ptr(fins), ptr(phs), ptr(ws), ws.numel(), _stream()))
    fin2, ph2 = fins.cpu().numpy(), phs.cpu().numpy()
    say("phantoms", ph1, ph2)
    assert np.array_equal(ph1, ph2) and np.all(ph2 > 0)
    np.testing.assert_allclose(fin2, fin1, rtol=1e-12, atol=1e-12)
    z2 = b.log_Z()
    np.testing.assert_allclose(z2, z1, rtol=1e-12, atol=1e-12)


def test_uniform_sampler_phase():
    """UniformSampler shells (samplers/uniform_samplers.py:42-85): every new sample satisfies
    the acceptance rule (logL > contour, or the cap of 100 likelihood evaluations)."""
    from jaxns_amd.engine import DeviceBatch, RunSpec
    say("uniform sampler shells")
    model, _, _, _ = mvn_problem(3, seed=2)
    spec = RunSpec(model=model, num_live_points=200, num_slices=15, max_samples=200 * 50, seed=5)
    b = DeviceBatch([spec], lanes_per_chain=4, waves_per_block=2)
    b.init_state()
    b.set_sampler(1)
    term = dict(max_samples=1e18)
    for it in range(6):
        b.collect_shell(term)
        st = b.states()[0]
        import torch
        torch.cuda.synchronize()
        ev = b.new_evals[:100].cpu().numpy()
        L = b.new_logL[:100].cpu().numpy()
        ok = (L > st.contour) | (L == st.contour) | (ev >= 100)
        assert ok.all(), it
        assert np.all(ev >= 1) and np.all(ev <= 100)
    b.set_sampler(0)


def test_run_queue_matches_batch():
    """engine.run_queue streams 11 mixture runs of uneven D and num_live_points through 4 slots (a slot
    takes the next run when its run terminates); every run's log Z, evaluation count, sample count and
    termination reason equal those of the same runs in one batch of the same launch shape, bit for bit."""
    from jaxns_amd.engine import DeviceBatch, RunSpec, run_queue
    say("run queue vs batch")
    specs = []
    rng = np.random.default_rng(4)
    for i in range(11):
        D = int(rng.integers(2, 9))
        N = int(rng.integers(60, 200))
        model, _, _, _ = gmm_problem(D, seed=100 + i)
        specs.append(RunSpec(model=model, num_live_points=N, num_slices=3 * D, k_phantom=i % 2,
                             max_samples=N * 150, seed=1000 + i))
    term = dict(dlogZ=math.log(1.0 + 1e-3), max_samples=1e18)
    q = run_queue(specs, 4, term, graph_iters=4, lanes_per_chain=2, waves_per_block=2)
    cap = dict(N=max(s.num_live_points for s in specs), max_samples=max(s.max_samples for s in specs),
               D=max(s.model.U_ndims for s in specs), num_slices=max(s.num_slices for s in specs),
               k_phantom=max(s.k_phantom for s in specs))
    b = DeviceBatch(specs, lanes_per_chain=2, waves_per_block=2, slot_capacity=cap)
    b.run(term, graph_iters=4)
    z = b.log_Z()
    sts = b.states()
    for i, r in enumerate(q):
        assert r["cap_hits"] == 0
        assert r["termination_reason"] == sts[i].termination_reason, i
        assert r["num_likelihood_evaluations"] == sts[i].num_likelihood_evaluations, i
        assert r["num_samples"] == sts[i].num_samples, i
        assert (r["log_Z_mean"], r["log_Z_uncert"]) == tuple(z[i]), i


def test_linreg_precision_at_scale():
    """The precision-at-scale regression (65 536 rows x 128 features, 10 % of the columns x 100,
    sigma = 0.1, w ~ N(0, I)): the device likelihood against the FP64 likelihood on the host at the same
    (FP32) U, on prior draws and on draws from the analytic posterior (logL ~ 5.8e4).  At the posterior
    the weights of the x100 columns are determined to ~4e-6, so w - w0 formed in FP32 (the FP32 ndtri,
    ~1e-7 relative in q, and the cancellation of w against w0 in column-scaled FP32 coordinates, ulp
    ~8e-6 at 100) moves them by a good part of their posterior width: both FP32-transform forms (BF16x3 = 3,
    plain FP32 = 1) are within 1e-5 relative on posterior draws (about 4e-6 measured; 2e-5 and 1e-6 on
    prior draws), while the refinement form (5: w - w0 from the FP64 transform and FP64 parameters, then
    the plain FP32 product) is within 5e-7 on both."""
    from scipy.special import ndtri
    torch = _torch()
    say("linreg 65536 x 128 precision")
    model, _, lik = linreg_problem(65536, 128, seed=0, sigma=0.1, tau=1.0, scaled_cols=0.1)
    rng = np.random.default_rng(1)
    B = 256
    U = np.empty((B, 128), np.float32)
    U[:B // 2] = rng.uniform(0.02, 0.98, size=(B // 2, 128))
    U[B // 2:] = _linreg_posterior_U(lik, 1.0, B - B // 2, rng).clip(1e-7, 1 - 1e-7)
    b = _batch(model, N=64, lpc=4, wpb=4)
    assert b.dps == 128
    Ud = torch.from_numpy(np.ascontiguousarray(U)).cuda()
    ref64 = lik.log_prob_numpy(ndtri(U.astype(np.float64)))
    qd = torch.empty_like(Ud)
    check(lib().hipns_ndtri_f32(ptr(Ud), Ud.numel(), ptr(qd), _stream()))
    ref32 = lik.log_prob_numpy(qd.cpu().numpy().astype(np.float64))
    P, Q = slice(0, B // 2), slice(B // 2, B)
    for terms, tol_prior, tol_post in ((3, 2e-5, 1e-5), (1, 1e-6, 1e-5), (5, 5e-7, 5e-7)):
        out = torch.empty(B, dtype=torch.float64, device="cuda")
        b.launch.lr_terms = terms
        check(lib().hipns_log_likelihood(ptr(Ud), B, ptr(out), ptr(b.runs_dev), b.runs_host, C.byref(b.bufs),
                                         C.byref(b.launch), _stream()))
        got = out.cpu().numpy()
        tot = np.abs(got - ref64) / np.abs(ref64)
        prod = np.abs(got - ref32) / np.abs(ref32)
        say(f"terms {terms}: against FP64 transform prior {tot[P].max():.2e} posterior {tot[Q].max():.2e} "
            f"(mean logL {ref64[Q].mean():.4g}); against the device's FP32 q prior {prod[P].max():.2e} "
            f"posterior {prod[Q].max():.2e}")
        assert tot[P].max() < tol_prior and tot[Q].max() < tol_post
