"""Statistical oracle and exact identities for the Philox-mode GPU engine.

* The reference's MC tests (mc.rs:370-455) at the same parameters, in FP64
  and FP32-path mode.
* GPU vs CPU oracle (restatement of mc.rs) at equal path counts: within 3
  combined standard errors.
* Identities that hold exactly: martingale and Asian-mean closed forms
  (within SE), knock-out at t=0, barrier at infinity == European,
  mc_bates(lambda=0) == mc_heston, determinism.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REF = dict(v0=0.04, kappa=2.0, theta=0.04, sigma=0.3, rho=-0.7)   # mc.rs:351-353
BASE = dict(v0=0.04, kappa=1.5, theta=0.04, sigma=0.5, rho=-0.7)  # the MC workloads' parameters
PRECISIONS = [0, 1]


def small_cfg(ope, scheme=0):
    return ope.McConfig(n_paths=60_000, n_steps=100, seed=7, antithetic=True, scheme=scheme)  # mc.rs:355-357


def run(ope, precision, payoff, cfg, hp=REF, jumps=None, spot=100.0, expiry=1.0, rate=0.05, q=0.0):
    print(f"gpu price precision={precision} payoff={payoff} cfg={cfg} jumps={jumps}", flush=True)
    return ope.mc_price_ex(spot, expiry, rate, q, ope.HestonParams(**hp), payoff, cfg, jumps,
                           ope.McOptions(precision))


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("scheme", [0, 1])
def test_mc_heston_matches_analytic_european(ope, ora, precision, scheme):
    analytic = ora.heston_price(100, 100, 1, 0.05, 0, ora.heston(**REF))
    r = run(ope, precision, ope.Payoff.European(100.0), small_cfg(ope, scheme)).result
    assert abs(r.price - analytic) / r.std_error < 4


@pytest.mark.parametrize("precision", PRECISIONS)
def test_mc_bates_matches_analytic_european(ope, ora, precision):
    analytic = ora.bates_price(100, 100, 1, 0.05, 0, ora.heston(**REF), 0.5, -0.1, 0.15)
    r = run(ope, precision, ope.Payoff.European(100.0), small_cfg(ope), jumps=(0.5, -0.1, 0.15)).result
    assert abs(r.price - analytic) / r.std_error < 4


@pytest.mark.parametrize("precision", PRECISIONS)
def test_asian_and_barrier_orderings(ope, precision):
    euro = run(ope, precision, ope.Payoff.European(100.0), small_cfg(ope)).result.price
    asian = run(ope, precision, ope.Payoff.AsianArithmetic(100.0), small_cfg(ope)).result.price
    uao = run(ope, precision, ope.Payoff.UpAndOut(100.0, 130.0), small_cfg(ope)).result.price
    assert asian < euro          # mc.rs:434-441
    assert 0.0 <= uao < euro     # mc.rs:446-455


@pytest.mark.parametrize("precision", PRECISIONS)
def test_qe_reduces_bias_in_feller_violating_regime(ope, ora, precision):
    hp = dict(v0=0.04, kappa=5.0, theta=0.04, sigma=1.2, rho=-0.7)   # mc.rs:370-391
    analytic = ora.heston_price(100, 100, 1, 0.05, 0, ora.heston(**hp))
    e = run(ope, precision, ope.Payoff.European(100.0),
            ope.McConfig(n_paths=200_000, n_steps=8, seed=42, scheme=0), hp=hp).result
    q = run(ope, precision, ope.Payoff.European(100.0),
            ope.McConfig(n_paths=200_000, n_steps=8, seed=42, scheme=1), hp=hp).result
    euler_err, qe_err = abs(e.price - analytic), abs(q.price - analytic)
    assert qe_err < euler_err / 5
    assert qe_err < 10 * max(q.std_error, 0.05)


STAT_CASES = [
    # (name, scheme, payoff args, jumps, n_paths, n_steps, rate, q)
    ("heston_euler_euro_100k_x_252", 0, ("E", 100.0), None, 100_000, 252, 0.03, 0.0),
    ("qe_asian", 1, ("A", 100.0), None, 100_000, 365, 0.03, 0.0),
    ("qe_uao", 1, ("U", 100.0, 130.0), None, 100_000, 365, 0.03, 0.0),
    ("bates_euler_euro_put", 0, ("P", 95.0), (0.5, -0.1, 0.15), 100_000, 252, 0.03, 0.01),
    ("bates_qe_uao", 1, ("U", 100.0, 120.0), (1.0, -0.05, 0.2), 60_000, 100, 0.03, 0.0),
]


def _payoff(ope, spec):
    if spec[0] == "E":
        return ope.Payoff.European(spec[1])
    if spec[0] == "P":
        return ope.Payoff.European(spec[1], ope.OptionType.Put)
    if spec[0] == "A":
        return ope.Payoff.AsianArithmetic(spec[1])
    return ope.Payoff.UpAndOut(spec[1], spec[2])


def _ora_payoff(ora, spec):
    if spec[0] == "E":
        return ora.payoff(ora.EUROPEAN, spec[1])
    if spec[0] == "P":
        return ora.payoff(ora.EUROPEAN, spec[1], ora.PUT)
    if spec[0] == "A":
        return ora.payoff(ora.ASIAN, spec[1])
    return ora.payoff(ora.UP_AND_OUT, spec[1], barrier=spec[2])


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("case", STAT_CASES, ids=[c[0] for c in STAT_CASES])
def test_gpu_matches_cpu_oracle_within_3_se(ope, ora, precision, case):
    name, scheme, pspec, jumps, n_paths, n_steps, rate, q = case
    cfg = ope.McConfig(n_paths=n_paths, n_steps=n_steps, seed=42, scheme=scheme)
    g = run(ope, precision, _payoff(ope, pspec), cfg, hp=BASE, jumps=jumps, rate=rate, q=q).result
    jp = ora.jumps(*jumps) if jumps else None
    c = ora.mc_price(100.0, 1.0, rate, q, ora.heston(**BASE), _ora_payoff(ora, pspec), n_paths, n_steps, 42,
                     True, scheme, jp, n_threads=8)
    assert abs(g.price - c[0]) < 3 * math.hypot(g.std_error, c[1]), (g, c[:2])
    assert 0.8 < g.std_error / c[1] < 1.25


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("scheme", [0, 1])
@pytest.mark.parametrize("jumps", [None, (0.5, -0.1, 0.15)])
def test_discounted_spot_is_a_martingale(ope, precision, scheme, jumps):
    # E[e^{-rT} S_T] = S0 e^{-qT}: pins the drift, including the jump compensator (mc.rs:155-157)
    cfg = ope.McConfig(n_paths=400_000, n_steps=50, seed=3, scheme=scheme)
    r = run(ope, precision, ope.Payoff.European(0.0), cfg, hp=BASE, jumps=jumps, rate=0.03, q=0.02).result
    assert abs(r.price - 100.0 * math.exp(-0.02)) < 4 * r.std_error


@pytest.mark.parametrize("precision", PRECISIONS)
def test_asian_mean_closed_form(ope, precision):
    # E[sum_{i=1..N} S_{t_i} / N] = S0 mean_i e^{(r-q) t_i}, S0 excluded (mc.rs:88, 271)
    n = 12
    cfg = ope.McConfig(n_paths=400_000, n_steps=n, seed=5)
    r = run(ope, precision, ope.Payoff.AsianArithmetic(0.0), cfg, hp=BASE, rate=0.03, q=0.0).result
    want = math.exp(-0.03) * 100.0 * np.mean([math.exp(0.03 * (i + 1) / n) for i in range(n)])
    assert abs(r.price - want) < 4 * r.std_error


@pytest.mark.parametrize("precision", PRECISIONS)
def test_barrier_at_or_below_spot_pays_discounted_rebate(ope, precision):
    cfg = ope.McConfig(n_paths=10_000, n_steps=20, seed=1)
    r = run(ope, precision, ope.Payoff.UpAndOut(100.0, 100.0, 3.0), cfg, hp=BASE, rate=0.05).result
    # every path pays the rebate discounted from T (mc.rs:218, 276-277); only
    # the FP64 summation rounding separates the mean from it
    assert abs(r.price - math.exp(-0.05) * 3.0) < 1e-13
    assert r.std_error < 1e-6


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("scheme", [0, 1])
@pytest.mark.parametrize("lam", [0.0, 0.5, 30.0])
def test_barrier_at_infinity_is_european(ope, precision, scheme, lam):
    # pathwise identical payoffs.  FP64 without jumps: bit-identical.  With
    # jumps the European kernels draw the jump sizes after the last step (from
    # three per-lane slots, inline beyond three) and add their sum once, while
    # the barrier kernel adds each jump in its step; in FP32-path mode the
    # European kernel also sums log-price increments in 8-step FP32 blocks.
    # So only rounding differs.  lam = 30 over 30 steps jumps at most steps,
    # overflowing the slots.
    cfg = ope.McConfig(n_paths=50_000, n_steps=30, seed=9, scheme=scheme)
    jumps = (lam, -0.1, 0.15) if lam > 0 else None
    e = run(ope, precision, ope.Payoff.European(100.0), cfg, hp=BASE, jumps=jumps).result
    u = run(ope, precision, ope.Payoff.UpAndOut(100.0, 1e300), cfg, hp=BASE, jumps=jumps).result
    if precision == 0 and lam == 0:
        assert e == u
    elif precision == 0:
        assert abs(e.price - u.price) < 1e-9 * e.std_error
    else:
        assert abs(e.price - u.price) < 1e-3 * e.std_error


@pytest.mark.parametrize("precision", PRECISIONS)
def test_bates_zero_intensity_is_heston_bitwise(ope, precision):
    cfg = ope.McConfig(n_paths=50_000, n_steps=30, seed=9)
    h = run(ope, precision, ope.Payoff.European(100.0), cfg, hp=BASE).result
    b = run(ope, precision, ope.Payoff.European(100.0), cfg, hp=BASE, jumps=(0.0, -0.1, 0.15)).result
    assert h == b


@pytest.mark.parametrize("precision", PRECISIONS)
def test_deterministic_and_seed_dependent(ope, precision):
    cfg = ope.McConfig(n_paths=2_200_000, n_steps=20, seed=11, scheme=1)   # 4 pairs per thread
    a = run(ope, precision, ope.Payoff.AsianArithmetic(100.0), cfg, hp=BASE, jumps=(0.5, -0.1, 0.15))
    b = run(ope, precision, ope.Payoff.AsianArithmetic(100.0), cfg, hp=BASE, jumps=(0.5, -0.1, 0.15))
    assert a.result == b.result
    cfg.seed = 12
    c = run(ope, precision, ope.Payoff.AsianArithmetic(100.0), cfg, hp=BASE, jumps=(0.5, -0.1, 0.15))
    assert c.result.price != a.result.price
    assert abs(c.result.price - a.result.price) < 4 * math.hypot(a.result.std_error, c.result.std_error)


def test_reference_api_shapes(ope):
    # mc_heston / mc_bates with the reference's McConfig::default()
    hp = ope.HestonParams(**REF)
    cfg = ope.McConfig()
    r1 = ope.mc_heston(100.0, 1.0, 0.05, 0.0, hp, ope.Payoff.AsianArithmetic(100.0), cfg)
    r2 = ope.mc_bates(100.0, 1.0, 0.05, 0.0, hp, 0.5, -0.10, 0.15, ope.Payoff.UpAndOut(100.0, 130.0, 0.0), cfg)
    assert 0 < r1.price < 10 and r1.std_error > 0
    assert 0 < r2.price < 10 and r2.std_error > 0


@pytest.mark.parametrize("scheme", [0, 1])
def test_long_paths_and_ragged_counts(ope, scheme):
    # 4096 steps (dt = 2.4e-4): the draw loops' blocks of 8 steps, the
    # 8-step FP32 partial sums and the FP64 fold over a long path; an odd path
    # count just above the 4-pairs-per-thread threshold (2^22 pairs) drops
    # its last path (mc.rs:160-161) and ends in a ragged tile.  FP32 and FP64
    # see the same draws, so they agree to a small fraction of the SE.
    hp = ope.HestonParams(**BASE)
    for n_paths, n_steps in ((20_001, 4096), (2 * (1 << 22) + 2 * 37 + 1, 13)):
        cfg = ope.McConfig(n_paths=n_paths, n_steps=n_steps, seed=77, scheme=scheme)
        for pay in (ope.Payoff.European(100.0), ope.Payoff.AsianArithmetic(100.0)):
            r64 = run(ope, 0, pay, cfg, hp=BASE, jumps=(0.5, -0.1, 0.15))
            r32 = run(ope, 1, pay, cfg, hp=BASE, jumps=(0.5, -0.1, 0.15))
            assert r64.stats.count == (n_paths // 2) * 2
            assert r64.stats.path_steps == (n_paths // 2) * 2 * n_steps
            assert abs(r32.result.price - r64.result.price) <= 0.1 * r64.result.std_error, (n_paths, n_steps, pay)


@pytest.mark.parametrize("scheme", [0, 1])
@pytest.mark.parametrize("jumps", [None, (0.5, -0.1, 0.15)])
def test_fp32_without_sibling_tracks_fp64(ope, scheme, jumps):
    # antithetic = false: one path per Philox pair (the FP32 QE loop is a
    # separate compiled case without the sibling); the FP64 kernel draws the
    # same words, so the two agree to a small fraction of the SE, and the
    # path count is n_paths itself, odd counts included
    pays = (ope.Payoff.European(100.0), ope.Payoff.AsianArithmetic(100.0), ope.Payoff.UpAndOut(100.0, 130.0))
    for n_paths, n_steps in ((40_001, 52), (300_000, 9)):
        cfg = ope.McConfig(n_paths=n_paths, n_steps=n_steps, seed=31, antithetic=False, scheme=scheme)
        for pay in pays:
            r64 = run(ope, 0, pay, cfg, hp=BASE, jumps=jumps)
            r32 = run(ope, 1, pay, cfg, hp=BASE, jumps=jumps)
            assert r32.stats.count == r64.stats.count == n_paths
            assert abs(r32.result.price - r64.result.price) <= 0.1 * r64.result.std_error, (n_paths, pay)


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("scheme", [0, 1])
def test_zero_variance_is_deterministic(ope, precision, scheme):
    # v0 = theta = 0: v stays 0 in both schemes (Euler: max(0 + 0, 0); QE:
    # m = s^2 = 0, the quadratic branch's a = m / (1 + b^2) = 0), so every
    # path is the forward S0 e^{(r - q) t} and the prices have closed forms
    # with SE 0: the European call e^{-rT} (F_T - K), the Asian call the
    # discounted mean of the forwards over the N monitoring dates minus K,
    # and the up-and-out call knocks out iff some forward reaches B
    s0, t, r, q, n = 100.0, 1.0, 0.05, 0.01, 64
    hp = dict(v0=0.0, kappa=1.5, theta=0.0, sigma=0.5, rho=-0.7)
    cfg = ope.McConfig(n_paths=4_000, n_steps=n, seed=3, scheme=scheme)
    fwd = s0 * np.exp((r - q) * t * np.arange(1, n + 1) / n)
    disc = math.exp(-r * t)
    tol = 1e-12 * s0 if precision == 0 else 2e-5 * s0
    cases = [(ope.Payoff.European(90.0), disc * (fwd[-1] - 90.0)),
             (ope.Payoff.AsianArithmetic(90.0), disc * (fwd.mean() - 90.0)),
             (ope.Payoff.UpAndOut(90.0, 103.0, 1.5), disc * 1.5),           # forward crosses 103 before T
             (ope.Payoff.UpAndOut(90.0, 110.0, 1.5), disc * (fwd[-1] - 90.0))]
    for pay, want in cases:
        res = run(ope, precision, pay, cfg, hp=hp, spot=s0, expiry=t, rate=r, q=q).result
        assert abs(res.price - want) <= tol, (pay, res.price, want)
        assert res.std_error <= 1e-6 * s0   # zero up to the rounding of sum and sum of squares


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("jumps", [None, (0.5, -0.1, 0.15)])
def test_constant_variance_is_black_scholes_or_merton(ope, ora, precision, jumps):
    # sigma = 0 with v0 = theta: Euler keeps v = theta, so ln S is exact GBM
    # (plus compound-Poisson jumps): European prices are Black-Scholes, or the
    # Merton (1976) series with jumps, within the MC error, for every strike
    hp = dict(v0=0.04, kappa=1.5, theta=0.04, sigma=0.0, rho=-0.7)
    cfg = ope.McConfig(n_paths=400_000, n_steps=50, seed=19, scheme=0)
    for k in (80.0, 100.0, 120.0):
        res = run(ope, precision, ope.Payoff.European(k), cfg, hp=hp, jumps=jumps, rate=0.03).result
        if jumps is None:
            want = ora.bsm_price(100.0, k, 1.0, 0.03, 0.0, 0.2)
        else:
            want = ora.merton_price(100.0, k, 1.0, 0.03, 0.0, 0.04, *jumps)
        assert abs(res.price - want) < 4 * res.std_error, (k, res.price, want, res.std_error)


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("scheme", [0, 1])
def test_one_step_asian_is_european_pathwise(ope, precision, scheme):
    # with one monitoring date the Asian average is S_T: same paths, same
    # payoffs up to rounding (FP64: the European path adds its jump after the
    # loop, (ln S0 + dy) + jump, the Asian one in the step, ln S0 + (dy +
    # jump), and each takes its own exponential, so S_T may differ in the
    # last bit; FP32 path: the Asian sum rounds S_T once more in FP32)
    cfg = ope.McConfig(n_paths=100_000, n_steps=1, seed=23, scheme=scheme)
    e = run(ope, precision, ope.Payoff.European(100.0), cfg, hp=BASE, jumps=(0.5, -0.1, 0.15)).result
    a = run(ope, precision, ope.Payoff.AsianArithmetic(100.0), cfg, hp=BASE, jumps=(0.5, -0.1, 0.15)).result
    if precision == 0:
        assert abs(a.price - e.price) <= 1e-13 * e.price and abs(a.std_error - e.std_error) <= 1e-12 * e.std_error
    else:
        # each S_T within a few FP32 roundings (1e-7 relative) of the other
        assert abs(a.price - e.price) <= 1e-5 * e.price and abs(a.std_error - e.std_error) <= 1e-4 * e.std_error


EDGE = [
    # (id, Heston overrides, expiry, rate, dividend yield, scheme, jumps)
    ("rho_minus_1", dict(rho=-1.0), 1.0, 0.03, 0.01, 0, None),
    ("rho_minus_1_qe_bates", dict(rho=-1.0), 1.0, 0.03, 0.01, 1, (0.5, -0.1, 0.15)),
    ("rho_plus_1_qe", dict(rho=1.0), 1.0, 0.03, 0.01, 1, None),
    ("rho_plus_1_bates", dict(rho=1.0), 1.0, 0.03, 0.01, 0, (0.5, -0.1, 0.15)),
    ("rho_zero_qe", dict(rho=0.0), 1.0, 0.03, 0.0, 1, None),
    ("negative_rate", {}, 2.0, -0.01, 0.05, 0, (0.5, -0.1, 0.15)),
    ("short_expiry_qe", {}, 1e-3, 0.03, 0.0, 1, (0.5, -0.1, 0.15)),
    ("fast_mean_reversion_qe", dict(kappa=50.0), 1.0, 0.03, 0.0, 1, None),
]


@pytest.mark.parametrize("case", EDGE, ids=[c[0] for c in EDGE])
def test_fp32_tracks_fp64_at_edge_parameters(ope, case):
    # the FP32 path mixes rho into the Box-Muller cosine through acos(rho)
    # (0 and pi/2 .. pi at the ends); |rho| = 1, rho = 0, a negative rate, a
    # 1e-3-year expiry and kappa dt = 12.5 stay within 0.1 SE of the FP64
    # kernel on the same Philox draws, for every payoff
    _, over, expiry, rate, q, scheme, jumps = case
    hp = dict(BASE, **over)
    cfg = ope.McConfig(n_paths=200_000, n_steps=3 if expiry < 0.01 else 24, seed=19, scheme=scheme)
    for pay in (ope.Payoff.European(100.0), ope.Payoff.European(95.0, ope.OptionType.Put),
                ope.Payoff.AsianArithmetic(100.0), ope.Payoff.UpAndOut(100.0, 130.0, 1.0)):
        r64 = run(ope, 0, pay, cfg, hp=hp, jumps=jumps, expiry=expiry, rate=rate, q=q).result
        r32 = run(ope, 1, pay, cfg, hp=hp, jumps=jumps, expiry=expiry, rate=rate, q=q).result
        assert math.isfinite(r32.price) and math.isfinite(r32.std_error), (r32, r64)
        # (the 95 put at the 1e-3-year expiry is ~8 sd out of the money: both 0)
        assert abs(r32.price - r64.price) <= 0.1 * r64.std_error + 1e-12, (case[0], pay, r32, r64)


@pytest.mark.parametrize("kind", ["asian", "uao"])
def test_gpu_matches_cpu_oracle_at_full_size_qe_1m_x_365(ope, ora, kind):
    # The headline configuration at its full size (Heston QE, 1M paths x 365
    # steps, arithmetic Asian call and up-and-out call B=130), both
    # precisions against the CPU oracle's mc_price at the same path count:
    # independent draws of the same scheme, so within 3 combined SE.
    import os

    hp = dict(v0=0.04, kappa=1.5, theta=0.04, sigma=0.5, rho=-0.7)
    spec = ("A", 100.0) if kind == "asian" else ("U", 100.0, 130.0)
    threads = max(1, min(16, os.cpu_count() or 1))
    print(f"oracle QE {kind} 1M x 365 on {threads} threads", flush=True)
    c = ora.mc_price(100.0, 1.0, 0.03, 0.0, ora.heston(**hp), _ora_payoff(ora, spec), 1_000_000, 365, 42,
                     True, 1, None, n_threads=threads)
    cfg = ope.McConfig(n_paths=1_000_000, n_steps=365, seed=42, scheme=1)
    for precision in PRECISIONS:
        g = run(ope, precision, _payoff(ope, spec), cfg, hp=hp, rate=0.03).result
        print(f"precision={precision} gpu={g} oracle=({c[0]}, {c[1]})", flush=True)
        assert abs(g.price - c[0]) < 3 * math.hypot(g.std_error, c[1]), (precision, g, c[:2])
        assert 0.97 < g.std_error / c[1] < 1.03
