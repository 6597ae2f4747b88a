"""GPU parity of the per-shell kernels against the oracle, each called alone through the C ABI on a
live set written by the test: replacement order (hipns_shell_replace), the termination register
(hipns_shell_register), the append of the discarded shell and of the final live set
(hipns_shell_discard, hipns_consume_live, including an append that runs past the end of the sample
collection), the seed choice and the slice directions of the fused sampler."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as O
from jaxns_amd import _lib
from jaxns_amd._lib import RunState, check, lib, ptr

from problems import mvn_problem

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def say(*a):
    print(*a, flush=True)


def _batch(Ns, D=4, max_samples=None, k_phantom=0):
    from jaxns_amd.engine import DeviceBatch, RunSpec
    model, _, _, _ = mvn_problem(D, seed=1)
    specs = [RunSpec(model=model, num_live_points=N, num_slices=3 * D, k_phantom=k_phantom,
                     max_samples=max_samples or N * 20, seed=i) for i, N in enumerate(Ns)]
    return DeviceBatch(specs, lanes_per_chain=2, waves_per_block=2)


def _put_states(b, sts):
    torch = _torch()
    raw = b"".join(bytes(s) for s in sts)
    b.states_dev.copy_(torch.frombuffer(bytearray(raw), dtype=torch.uint8))


def _args(b):
    return (ptr(b.runs_dev), ptr(b.states_dev), b.R, b.runs_host, C.byref(b.bufs))


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.int64)


def _keys(rng, n, scale=8):
    """logL keys with many ties (values on a coarse grid), some -inf and a few NaN."""
    k = np.round(rng.normal(size=n) * scale) / 4.0
    k[rng.random(n) < 0.05] = -np.inf
    k[rng.random(n) < 0.01] = np.nan
    return k


# --------------------------------------------------------------------------- replacement order
@pytest.mark.parametrize("Ns,cur", [((64,), 0), ((64, 1000, 8192), 1), ((301, 2, 4096), 0)])
def test_shell_replace_matches_stable_argsort(Ns, cur):
    """live[:S] = new samples, then the stable argsort of the whole live set (sharded_static.py:262-275):
    the device's new live set equals numpy.argsort(kind='stable') of [new, survivors] applied to every
    field, bit for bit, with ties between and within the two parts, -inf and NaN keys."""
    torch = _torch()
    say(f"shell_replace N={Ns} cur={cur}")
    rng = np.random.default_rng(sum(Ns) + cur)
    b = _batch(Ns)
    dps = b.dps
    sts = b.states()
    exp = []
    for r, N in enumerate(Ns):
        d = b.runs_host[r]
        S, lo, co = int(d.shell), int(d.live_off), int(d.chain_off)
        live = np.sort(_keys(rng, N))  # NaN last, as the sorted live set holds them
        new = _keys(rng, S)
        new[: S // 3] = rng.choice(live, S // 3)  # exact ties with survivors
        U_live = rng.random((N, dps)).astype(np.float32)
        U_new = rng.random((S, dps)).astype(np.float32)
        cons = rng.normal(size=N)
        send = rng.integers(-1, 10_000, N).astype(np.int32)
        ev_live = rng.integers(0, 500, N).astype(np.int32)
        ev_new = rng.integers(0, 500, S).astype(np.int32)
        contour, sender = float(rng.normal()), int(rng.integers(0, 10_000))
        b.live_logL[cur][lo:lo + N] = torch.from_numpy(live)
        b.live_U[cur][lo:lo + N] = torch.from_numpy(U_live)
        b.live_cons[cur][lo:lo + N] = torch.from_numpy(cons)
        b.live_sender[cur][lo:lo + N] = torch.from_numpy(send)
        b.live_evals[cur][lo:lo + N] = torch.from_numpy(ev_live)
        b.new_logL[co:co + S] = torch.from_numpy(new)
        b.new_U[co:co + S] = torch.from_numpy(U_new)
        b.new_evals[co:co + S] = torch.from_numpy(ev_new)
        sts[r].buf, sts[r].done, sts[r].contour, sts[r].sender_node = cur, 0, contour, sender
        keys = np.concatenate([new, live[S:]])
        perm = np.argsort(keys, kind="stable")
        exp.append(dict(logL=keys[perm], U=np.concatenate([U_new, U_live[S:]])[perm],
                        cons=np.concatenate([np.full(S, contour), cons[S:]])[perm],
                        send=np.concatenate([np.full(S, sender, np.int32), send[S:]])[perm],
                        ev=np.concatenate([ev_new, ev_live[S:]])[perm]))
    _put_states(b, sts)
    check(lib().hipns_shell_replace(*_args(b), _lib.stream_ptr()), "hipns_shell_replace")
    torch.cuda.synchronize()
    nxt = 1 - cur
    for r, N in enumerate(Ns):
        lo = int(b.runs_host[r].live_off)
        e = exp[r]
        assert np.array_equal(_bits(b.live_logL[nxt][lo:lo + N].cpu().numpy()), _bits(e["logL"])), r
        assert np.array_equal(b.live_U[nxt][lo:lo + N].cpu().numpy(), e["U"]), r
        assert np.array_equal(_bits(b.live_cons[nxt][lo:lo + N].cpu().numpy()), _bits(e["cons"])), r
        assert np.array_equal(b.live_sender[nxt][lo:lo + N].cpu().numpy(), e["send"]), r
        assert np.array_equal(b.live_evals[nxt][lo:lo + N].cpu().numpy(), e["ev"]), r


# --------------------------------------------------------------------------- termination register
def _oracle_register(prev8, disc, live, live_evals, new_evals, N, contour, prev_nle, prev_peak, num_samples):
    """_collect_shell's register update (sharded_static.py:281-323) with the sequential oracle recurrence."""
    ev = np.array(prev8, dtype=np.float64)
    for L in disc:
        ev = O.evidence_update(ev, N, L)
    evr = ev.copy()
    for i, L in enumerate(live):
        evr = O.evidence_update(evr, N - i, L)
    S = disc.size
    reg = O.Register()
    reg.num_samples_used = num_samples
    for i in range(8):
        reg.evidence_calc[i] = ev[i]
        reg.evidence_calc_with_remaining[i] = evr[i]
    reg.num_likelihood_evaluations = prev_nle + int(np.sum(new_evals))
    reg.log_L_contour = contour
    reg.efficiency = N / float(np.sum(live_evals))
    reg.plateau = int(np.all(live == live[0]))
    reg.no_seed_points = int(live[S - 1] >= live[N - 1])
    reg.absolute_spread = abs(live[-1] - live[0])
    reg.relative_spread = 2.0 * reg.absolute_spread / abs(live[0] + live[-1])
    pxl = ev[1] + ev[0]
    reg.peak_log_XL = np.nan if (np.isnan(prev_peak) or np.isnan(pxl)) else max(prev_peak, pxl)
    return reg


TERMS = [dict(dlogZ=math.log(1.0 + 1e-3), max_samples=1e9),
         dict(dlogZ=5.0, max_samples=1e9),
         dict(ess=50.0, evidence_uncert=0.5, efficiency_threshold=0.2),
         dict(max_samples=600.0, max_num_likelihood_evaluations=1e5, log_L_contour=1.0),
         dict(rtol=0.5, atol=3.0, peak_XL_frac=0.5)]


@pytest.mark.parametrize("Ns,plateau", [((64,), False), ((64, 500, 2048), False), ((128,), True)])
def test_shell_register_matches_sequential_oracle(Ns, plateau):
    """hipns_shell_register (map composition over the discarded shell and the remaining live points)
    against the sequential _update_evidence_calc_op of the oracle: both EvidenceCalculations to 1e-11,
    and efficiency, plateau, spreads, no_seed_points, peak_log_XL, counters exactly; the termination
    bits equal determine_termination of the oracle on the same register, for several conditions."""
    torch = _torch()
    say(f"shell_register N={Ns} plateau={plateau}")
    rng = np.random.default_rng(len(Ns) * 7 + plateau)
    for ti, term in enumerate(TERMS):
        b = _batch(Ns)
        sts = b.states()
        exp = []
        for r, N in enumerate(Ns):
            d = b.runs_host[r]
            S, lo, co = int(d.shell), int(d.live_off), int(d.chain_off)
            # a register after three earlier shells
            prev = O.evidence_init()
            Lp = np.sort(rng.normal(size=3 * S)) * 3 - 20
            for L in Lp:
                prev = O.evidence_update(prev, N, L)
            disc = np.sort(Lp[-1] + rng.exponential(size=S))
            live = np.sort(disc[-1] + rng.exponential(size=N) * (0.0 if plateau else 2.0))
            if not plateau:
                live[rng.integers(0, N, 3)] = disc[-1]  # points on the contour
                live = np.sort(live)
            live_ev = rng.integers(1, 60, N).astype(np.int32)
            new_ev = rng.integers(1, 60, S).astype(np.int32)
            cur = int(rng.integers(0, 2))
            b.live_logL[cur][lo:lo + N] = torch.from_numpy(np.concatenate([disc, rng.normal(size=N - S)]))
            b.live_logL[1 - cur][lo:lo + N] = torch.from_numpy(live)
            b.live_evals[1 - cur][lo:lo + N] = torch.from_numpy(live_ev)
            b.new_evals[co:co + S] = torch.from_numpy(new_ev)
            st = sts[r]
            st.buf, st.done, st.contour = cur, 0, float(disc[-1])
            for i in range(8):
                st.evidence_calc[i] = prev[i]
            st.num_samples = 3 * S
            st.next_idx = 3 * S
            st.iteration = 3
            st.num_likelihood_evaluations = int(rng.integers(0, 10_000))
            st.peak_log_XL = float(prev[1] + prev[0] - 0.5)
            reg = _oracle_register(prev, disc, live, live_ev, new_ev, N, float(disc[-1]), st.num_likelihood_evaluations,
                                   st.peak_log_XL, 4 * S)
            exp.append((reg, st.num_likelihood_evaluations))
        _put_states(b, sts)
        tc = _lib.make_term_cond(**term)
        check(lib().hipns_shell_register(*_args(b), C.byref(tc), _lib.stream_ptr()), "hipns_shell_register")
        got = b.states()
        for r, N in enumerate(Ns):
            g, (reg, _) = got[r], exp[r]
            S = int(b.runs_host[r].shell)
            np.testing.assert_allclose(np.array(g.evidence_calc[:]), np.array(reg.evidence_calc[:]), rtol=1e-11,
                                       atol=1e-11)
            np.testing.assert_allclose(np.array(g.evidence_calc_with_remaining[:]),
                                       np.array(reg.evidence_calc_with_remaining[:]), rtol=1e-11, atol=1e-11)
            assert g.efficiency == reg.efficiency
            assert g.plateau == reg.plateau == int(plateau)
            assert g.no_seed_points == reg.no_seed_points
            assert g.absolute_spread == reg.absolute_spread
            assert _bits([g.relative_spread])[0] == _bits([reg.relative_spread])[0]
            np.testing.assert_allclose(g.peak_log_XL, reg.peak_log_XL, rtol=1e-12)
            assert g.num_likelihood_evaluations == reg.num_likelihood_evaluations
            assert g.log_L_contour == reg.log_L_contour
            assert g.num_samples == g.num_samples_used == 4 * S
            assert g.next_idx == (4 * S) % int(b.runs_host[r].max_samples)
            assert g.iteration == 4
            assert g.buf == 1 - sts[r].buf
            # termination: the oracle's determine_termination on the oracle register, with max_samples
            # capped to leave room for one more shell (sharded_static.py:461-470)
            kw = dict(term)
            if "max_samples" in kw:
                kw["max_samples"] = min(kw["max_samples"], int(b.runs_host[r].max_samples) - S)
            done, why = O.determine_termination(kw, reg)
            assert (bool(g.done), int(g.termination_reason)) == (done, why), (ti, r, why, g.termination_reason)


def test_uniform_shell_register_adds_no_phantoms():
    """A UniformSampler shell keeps no phantom samples (num_phantom() = 0, samplers/uniform_samplers.py),
    so with k = 2 the register advances the collection by the shell alone, and the max_samples cap
    leaves room for one shell without phantoms (sharded_static.py:461-470): at 1800 samples of a
    1920-sample collection (S = 32) the run continues, where a cap of 1920 - 3 S = 1824 would stop it."""
    torch = _torch()
    say("uniform shell register with k=2")
    b = _batch((64,), k_phantom=2, max_samples=64 * 30)
    b.set_sampler(1)
    sts = b.states()
    S, N, ms = int(b.runs_host[0].shell), 64, int(b.runs_host[0].max_samples)
    assert (S, ms) == (32, 1920)
    b.live_logL[0][:N] = torch.from_numpy(np.sort(np.random.default_rng(0).normal(size=N)))
    b.live_logL[1][:N] = torch.from_numpy(np.sort(np.random.default_rng(1).normal(size=N)) + 3)
    b.live_evals[1][:N] = 2
    sts[0].buf, sts[0].num_samples, sts[0].next_idx, sts[0].done = 0, 1800, 1800, 0
    _put_states(b, sts)
    tc = _lib.make_term_cond(max_samples=1e9)
    check(lib().hipns_shell_register(*_args(b), C.byref(tc), _lib.stream_ptr()), "hipns_shell_register")
    g = b.states()[0]
    assert g.num_samples == 1800 + S and g.next_idx == 1800 + S
    assert g.done == 0 and g.termination_reason == 0
    # the same run's cap before a shell (hipns_check_termination): 1832 < 1920 - 32 keeps it running
    b.check_termination(dict(max_samples=1e9), set_no_seed=False)
    assert b.states()[0].done == 0


# --------------------------------------------------------------------------- append (discard / consume)
@pytest.mark.parametrize("next_idx", [0, 100, 150, 159])
def test_shell_discard_and_consume_placement(next_idx):
    """_add_samples_to_state through replace_index (sharded_static.py:40-85, internals/maps.py:15-25):
    the discarded shell lands at next_sample_idx, or, when it would run past the end of the collection,
    at the clamped start max_samples - S (lax.dynamic_update_slice); next_sample_idx then wraps modulo
    max_samples.  Contour, first seed index and sender node (sharded_static.py:250-261) and the final
    append of the live set (:834-838) follow the same rule."""
    torch = _torch()
    say(f"discard/consume next_idx={next_idx}")
    rng = np.random.default_rng(next_idx)
    b = _batch((64,), max_samples=160)
    ms, N, S, dps = int(b.runs_host[0].max_samples), 64, int(b.runs_host[0].shell), b.dps
    assert ms == 160
    cur = 1
    live = np.sort(np.round(rng.normal(size=N) * 4) / 2)
    U = rng.random((N, dps)).astype(np.float32)
    send = rng.integers(0, 1000, N).astype(np.int32)
    ev = rng.integers(1, 90, N).astype(np.int32)
    for t, v in ((b.live_logL[cur], live), (b.live_U[cur], U), (b.live_sender[cur], send), (b.live_evals[cur], ev)):
        t[:N] = torch.from_numpy(v)
    b.samp_logL.fill_(-7.0)
    b.samp_U.fill_(-1.0)
    b.samp_phantom.fill_(1)
    sts = b.states()
    sts[0].buf, sts[0].next_idx, sts[0].num_samples, sts[0].done = cur, next_idx, next_idx, 0
    _put_states(b, sts)
    check(lib().hipns_shell_discard(*_args(b), _lib.stream_ptr()), "hipns_shell_discard")
    g = b.states()[0]
    start = O.update_start(next_idx, S, ms)
    sl = slice(start, start + S)
    assert np.array_equal(b.samp_logL[sl].cpu().numpy(), live[:S])
    assert np.array_equal(b.samp_U[sl].cpu().numpy(), U[:S])
    assert np.array_equal(b.samp_sender[sl].cpu().numpy(), send[:S])
    assert np.array_equal(b.samp_evals[sl].cpu().numpy(), ev[:S])
    assert not b.samp_phantom[sl].cpu().numpy().any()
    rest = np.ones(ms, bool)
    rest[sl] = False
    assert np.all(b.samp_logL.cpu().numpy()[rest] == -7.0)
    contour = live[S - 1]
    assert g.contour == contour
    assert g.first_seed == int(np.searchsorted(live, contour, side="right"))
    assert g.sender_node == (next_idx + S) % ms - 1
    # final append of the live set
    b.samp_logL.fill_(-7.0)
    check(lib().hipns_consume_live(*_args(b), _lib.stream_ptr()), "hipns_consume_live")
    g = b.states()[0]
    start = O.update_start(next_idx, N, ms)
    assert np.array_equal(b.samp_logL[start:start + N].cpu().numpy(), live)
    assert np.array_equal(b.samp_U[start:start + N].cpu().numpy(), U)
    assert g.num_samples == next_idx + N and g.next_idx == (next_idx + N) % ms


# --------------------------------------------------------------------------- draws of the fused sampler
def _u01d(a, b):
    v = ((int(a) << 21) ^ (int(b) >> 11)) & ((1 << 53) - 1)
    return v / 9007199254740992.0


@pytest.mark.parametrize("N,first", [(2048, 700), (2048, 0), (2048, 2047), (300, 300), (5, 2)])
def test_seed_indices_match_oracle(N, first):
    """get_seed_point / resample_indicies (uni_slice_sampler.py:343-358, random.py:34-84): the device's
    closed form first + ceil(M (1 - u)) - 1 picks, for every chain's uniform, the index that the
    oracle's cumulative-logsumexp + searchsorted restatement picks on a live set whose entries
    [first, N) lie above the contour; the uniform is the 53-bit mapping of the chain's Philox block."""
    torch = _torch()
    say(f"seed indices N={N} first={first}")
    C_ = 3000
    k0, k1, it = 0x1234567, 0x89ABCDE, 77
    idx = torch.empty(C_, dtype=torch.int32, device="cuda")
    u = torch.empty(C_, dtype=torch.float64, device="cuda")
    check(lib().hipns_seed_indices(k0, k1, it, first, N, C_, ptr(idx), ptr(u), _lib.stream_ptr()))
    idx, u = idx.cpu().numpy(), u.cpu().numpy()
    rng = np.random.default_rng(N + first)
    live = np.sort(rng.normal(size=N))
    contour = live[first - 1] if first > 0 else -np.inf
    live[:first] = np.minimum(live[:first], contour)
    for c in range(0, C_, 7):
        w = O.philox4x32([c, it, 0xFFFFFFFF, 0], [k0, k1])
        assert u[c] == _u01d(w[0], w[1]), c
        assert idx[c] == O.seed_index(live, contour, u[c]), (c, idx[c])
    if first < N:
        # uniform over [first, N)
        assert idx.min() >= first and idx.max() <= N - 1
        if N - first > 1:
            h = np.bincount(idx - first, minlength=N - first)
            assert h.sum() == C_


def _oracle_direction(k0, k1, it, chain, si, D):
    """Box-Muller normals from two Philox blocks per 8 dimensions (radius uniforms from the first,
    angle uniforms from the second, 24 bits each), normalised."""
    if D == 1:
        return np.ones(1)
    z = np.zeros(((D + 7) // 8) * 8)
    for blk in range(z.size // 8):
        w = O.philox4x32([chain, it, si, (1 << 24) | blk], [k0, k1])
        v = O.philox4x32([chain, it, si, (4 << 24) | blk], [k0, k1])
        for q in range(4):
            u1 = ((int(w[q]) >> 8) + 1) / 16777216.0
            u2 = (int(v[q]) >> 8) / 16777216.0
            rr = math.sqrt(-2.0 * math.log(u1))
            z[8 * blk + 2 * q] = rr * math.cos(2 * math.pi * u2)
            z[8 * blk + 2 * q + 1] = rr * math.sin(2 * math.pi * u2)
    z = z[:D]
    return z / np.linalg.norm(z)


@pytest.mark.parametrize("D,lpc", [(1, 2), (3, 4), (10, 2), (32, 4), (33, 2), (70, 4), (128, 2)])
def test_slice_directions(D, lpc):
    """_sample_direction (uni_slice_sampler.py:23-38): unit vectors, equal to the FP64 restatement of
    the kernel's Box-Muller draws (24-bit uniforms) to 2e-5 per component (hardware log2 / sin / cos), and
    isotropic: n_1^2 follows Beta(1/2, (D-1)/2) (KS test over 24 576 directions)."""
    from scipy import stats
    torch = _torch()
    say(f"slice directions D={D} lpc={lpc}")
    C_, NS = 1024, 24
    k0, k1, it = 99, 12345, 3
    out = torch.zeros((C_, NS, D), dtype=torch.float32, device="cuda")
    check(lib().hipns_slice_directions(k0, k1, it, C_, NS, D, lpc, ptr(out), _lib.stream_ptr()))
    n = out.cpu().numpy().astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(n, axis=2), 1.0, atol=2e-6)
    for c, si in [(0, 0), (1, 5), (63, 23), (500, 11), (1023, 7)]:
        np.testing.assert_allclose(n[c, si], _oracle_direction(k0, k1, it, c, si, D), atol=2e-5)
    if D >= 2:
        x = n[:, :, 0].ravel() ** 2
        p = stats.kstest(x, stats.beta(0.5, (D - 1) / 2.0).cdf).pvalue
        say("KS p", p)
        assert p > 1e-4


# --------------------------------------------------------------------------- uniform draws (phase 1, init)
def _uniform_U(k0, k1, it, chain, attempt, D):
    """Model.sample_U of the fused kernel: dimension 4b + q from word q of the chain's Philox block
    (chain, iteration, attempt, UNIF | b), as (w >> 8) 2^-24 (exact in binary32)."""
    u = np.empty(D, np.float32)
    for blk in range((D + 3) // 4):
        w = O.philox4x32([chain, it, attempt, (3 << 24) | blk], [k0, k1])
        for q in range(4):
            if 4 * blk + q < D:
                u[4 * blk + q] = np.float32((int(w[q]) >> 8) * 2.0 ** -24)
    return u


def test_uniform_draws_match_oracle():
    """create_init_state's uniform draws (initialisation.py:20-84, uniform_sample.py:12-61) and one
    UniformSampler shell (uniform_samplers.py:42-85) against the same draws restated on the host:
    every initial live point is the chain's first draw (logL > -inf), and every chain of the shell
    accepts the first draw with logL >= contour (or the 100th), with the oracle's FP64 likelihood
    deciding; U bit for bit, evaluation counts exactly, logL to 1e-5."""
    torch = _torch()
    say("uniform draws vs oracle")
    from problems import mvn_oracle_spec
    D = 5
    model, _, mu, Sig = mvn_problem(D, seed=1)
    spec = mvn_oracle_spec(D, mu, Sig)
    b = _batch((64,), D=D)
    k0, k1 = int(b.runs_host[0].seed_lo), int(b.runs_host[0].seed_hi)
    b.init_state()
    st = b.states()[0]
    N, S = 64, int(b.runs_host[0].shell)
    live_U = b.live_U[st.buf][:N, :D].cpu().numpy()
    host = np.stack([_uniform_U(k0, k1, 0xFFFFFFFF, c, 0, D) for c in range(N)])
    key = lambda a: sorted(map(bytes, np.ascontiguousarray(a)))  # noqa: E731
    assert key(live_U) == key(host)
    assert np.all(b.live_evals[st.buf][:N].cpu().numpy() == 1)
    # one uniform-sampler shell from the initial live set
    b.set_sampler(1)
    args = _args(b)
    check(lib().hipns_shell_discard(*args, _lib.stream_ptr()), "hipns_shell_discard")
    check(lib().hipns_shell_sample(*args, C.byref(b.launch), _lib.stream_ptr()), "hipns_shell_sample")
    st = b.states()[0]
    contour, it = st.contour, int(st.iteration)
    U = b.new_U[:S, :D].cpu().numpy()
    L = b.new_logL[:S].cpu().numpy()
    ev = b.new_evals[:S].cpu().numpy()
    checked = 0
    for c in range(S):
        a, close = 0, False
        while True:
            u = _uniform_U(k0, k1, it, c, a, D)
            lo = float(spec.log_likelihood(u.astype(np.float64))[0])
            close |= abs(lo - contour) < 1e-4 * max(1.0, abs(contour))
            a += 1
            if lo >= contour or a >= 100:
                break
        if close:
            continue  # an FP32-vs-FP64 difference could flip this chain's decision
        checked += 1
        assert ev[c] == a, c
        assert np.array_equal(U[c], u), c
        assert abs(L[c] - lo) <= 1e-5 * max(1.0, abs(lo)), c
    say("chains checked", checked, "of", S)
    assert checked >= S - 2
    b.set_sampler(0)


def test_sampler_plugin_reproduces_a_shell():
    """The sampler plugin interface (samplers/abc.py:18-115; UniDimSliceSampler,
    uni_slice_sampler.py:276-441): pre_process of the live points, then get_samples with a run's seed,
    shell index and contour, gives that shell's new samples bit for bit (the engine's shell and the
    plugin run the same kernel on the same stream); every sample is above the contour, phantom samples
    count k per sample with no likelihood evaluations, and get_sample returns one sample."""
    import torch
    from jaxns_amd.samplers import EphemeralState, LivePointCollection, UniDimSliceSampler
    say("sampler plugin vs shell")
    D, N, k = 6, 128, 2
    model, _, _, _ = mvn_problem(D, seed=3)
    from jaxns_amd.engine import DeviceBatch, RunSpec
    spec = RunSpec(model=model, num_live_points=N, num_slices=3 * D, k_phantom=k, max_samples=N * 40, seed=77)
    b = DeviceBatch([spec], lanes_per_chain=2, waves_per_block=2)
    b.init_state()
    for _ in range(3):
        b.collect_shell(dict(max_samples=1e18))
    st = b.states()[0]
    cur = st.buf
    U = b.live_U[cur][:N, :D].double().cpu().numpy()
    L = b.live_logL[cur][:N].cpu().numpy()
    S = int(b.runs_host[0].shell)
    args = _args(b)
    check(lib().hipns_shell_discard(*args, _lib.stream_ptr()), "hipns_shell_discard")
    check(lib().hipns_shell_sample(*args, C.byref(b.launch), _lib.stream_ptr()), "hipns_shell_sample")
    st = b.states()[0]
    want_U = b.new_U[:S, :D].double().cpu().numpy()
    want_L = b.new_logL[:S].cpu().numpy()
    sampler = UniDimSliceSampler(model, num_slices=3 * D, num_phantom_save=k, midpoint_shrink=True)
    assert sampler.num_phantom() == k
    lp = LivePointCollection(sender_node_idx=np.zeros(N, np.int64), U_sample=U, log_L_constraint=np.zeros(N),
                             log_L=L, num_likelihood_evaluations=np.ones(N, np.int64))
    state = sampler.pre_process(EphemeralState(key=0, live_points_collection=lp))
    s, ph = sampler.get_samples((77, int(st.iteration)), st.contour, state, S)
    assert np.array_equal(s.U_sample, want_U)
    assert np.array_equal(s.log_L, want_L)
    assert np.all(s.log_L > st.contour)
    assert ph.log_L.shape == (S * k,) and np.all(ph.num_likelihood_evaluations == 0) and np.all(ph.log_L > st.contour)
    one, _ = sampler.get_sample(5, st.contour, state)
    assert one.U_sample.shape == (D,) and one.log_L > st.contour
