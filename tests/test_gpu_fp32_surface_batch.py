"""FP32-path mode against the FP64 GPU kernel on identical Philox streams
(|price_fp32 - price_fp64| <= 0.1 standard error), the shared-path surface
against single calls, and the one-launch uneven batch against single calls
(both bit-identical: the partial-sum order depends on the request alone)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BASE = dict(v0=0.04, kappa=1.5, theta=0.04, sigma=0.5, rho=-0.7)
JUMP = (0.5, -0.1, 0.15)


def both(ope, payoff, cfg, jumps=None, hp=BASE, expiry=1.0):
    out = []
    for precision in (0, 1):
        print(f"precision={precision} payoff={payoff} cfg={cfg} jumps={jumps}", flush=True)
        out.append(ope.mc_price_ex(100.0, expiry, 0.03, 0.0, ope.HestonParams(**hp), payoff, cfg, jumps,
                                   ope.McOptions(precision)).result)
    return out


@pytest.mark.parametrize("payoff", ["asian", "uao", "euro"])
def test_fp32_within_tenth_se_qe_1m_x_365(ope, payoff):
    # Heston QE, 1M paths x 365 steps (the Asian / up-and-out configuration)
    cfg = ope.McConfig(n_paths=1_000_000, n_steps=365, seed=42, scheme=1)
    p = {"asian": ope.Payoff.AsianArithmetic(100.0), "uao": ope.Payoff.UpAndOut(100.0, 130.0),
         "euro": ope.Payoff.European(100.0)}[payoff]
    f64, f32 = both(ope, p, cfg)
    print(f"fp64={f64} fp32={f32} err/se={abs(f32.price - f64.price) / f64.std_error:.4f}", flush=True)
    assert abs(f32.price - f64.price) <= 0.1 * f64.std_error
    assert abs(f32.std_error / f64.std_error - 1) < 1e-3


@pytest.mark.parametrize("scheme", [0, 1])
def test_fp32_within_tenth_se_bates(ope, scheme):
    cfg = ope.McConfig(n_paths=1_000_000, n_steps=252, seed=42, scheme=scheme)
    f64, f32 = both(ope, ope.Payoff.European(100.0), cfg, JUMP)
    assert abs(f32.price - f64.price) <= 0.1 * f64.std_error


def test_fp32_surface_within_tenth_se_full_size(ope, ora):
    # Bates, 4M paths x 252 steps, 41 strikes 60..140 at 8 expiries, FP32 vs
    # FP64; two of the entries also against the CPU oracle's mc_bates at the
    # same path count (independent draws: within 3 combined SE)
    import os

    strikes = np.arange(60.0, 140.0 + 1e-9, 2.0)
    expiries = [0.1, 0.25, 0.5, 0.75, 1.0, 1.25, 1.5, 2.0]
    cfg = ope.McConfig(n_paths=4_000_000, n_steps=252, seed=42)
    hp = ope.HestonParams(**BASE)
    p64, s64, st64, ms64 = ope.mc_surface(100.0, 0.03, 0.0, hp, JUMP, 0, expiries, strikes, cfg, 0)
    p32, s32, st32, ms32 = ope.mc_surface(100.0, 0.03, 0.0, hp, JUMP, 0, expiries, strikes, cfg, 1)
    err = np.abs(p32 - p64) / s64
    print(f"surface max |fp32-fp64|/se = {err.max():.4f}; fp64 {ms64:.2f} ms, fp32 {ms32:.2f} ms", flush=True)
    assert err.max() <= 0.1
    assert st64[0].path_steps == 4_000_000 * 252
    threads = max(1, min(16, os.cpu_count() or 1))
    for ie, ik in ((4, 20), (1, 15)):   # (T=1, K=100), (T=0.25, K=90)
        T, K = expiries[ie], float(strikes[ik])
        print(f"oracle Bates T={T} K={K} 4M x 252 on {threads} threads", flush=True)
        c = ora.mc_price(100.0, T, 0.03, 0.0, ora.heston(**BASE), ora.payoff(ora.EUROPEAN, K), 4_000_000, 252, 42,
                         True, 0, ora.jumps(*JUMP), n_threads=threads)
        for p, se in ((p64, s64), (p32, s32)):
            print(f"  gpu {p[ie, ik]} +- {se[ie, ik]}, oracle {c[0]} +- {c[1]}", flush=True)
            assert abs(p[ie, ik] - c[0]) < 3 * math.hypot(se[ie, ik], c[1])
            assert 0.97 < se[ie, ik] / c[1] < 1.03


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("n_paths", [30_000, 2_100_000])
def test_surface_equals_single_calls_bitwise(ope, precision, n_paths):
    strikes = [80.0, 95.0, 100.0, 117.5, 130.0]
    expiries = [0.25, 1.0, 1.5]
    cfg = ope.McConfig(n_paths=n_paths, n_steps=24, seed=5)
    hp = ope.HestonParams(**BASE)
    prices, ses, _, _ = ope.mc_surface(100.0, 0.03, 0.0, hp, JUMP, 1, expiries, strikes, cfg, precision)
    for e, t in enumerate(expiries):
        for k, K in enumerate(strikes):
            if n_paths > 1_000_000 and k not in (0, 4):
                continue
            r = ope.mc_price_ex(100.0, t, 0.03, 0.0, hp, ope.Payoff.European(K, ope.OptionType.Put), cfg, JUMP,
                                ope.McOptions(precision)).result
            assert (prices[e, k], ses[e, k]) == (r.price, r.std_error), (t, K)


@pytest.mark.parametrize("precision", [0, 1])
def test_direct_and_staged_launches_agree_bitwise(ope, precision):
    # One request per launch group with at most 16 strikes runs as a direct
    # launch (descriptor and strikes in the kernel arguments, the last block
    # reducing the slots); 17 strikes take the staged upload and the separate
    # reduction launch.  The 16 shared strikes agree bit for bit, and both
    # equal single calls (direct, one strike each).
    strikes = list(np.linspace(70.0, 130.0, 17))
    cfg = ope.McConfig(n_paths=300_000, n_steps=40, seed=9, scheme=1)
    hp = ope.HestonParams(**BASE)
    p16, s16, _, _ = ope.mc_surface(100.0, 0.03, 0.0, hp, JUMP, 0, [0.75], strikes[:16], cfg, precision)
    p17, s17, _, _ = ope.mc_surface(100.0, 0.03, 0.0, hp, JUMP, 0, [0.75], strikes, cfg, precision)
    assert np.array_equal(p16[0], p17[0, :16]) and np.array_equal(s16[0], s17[0, :16])
    for k in (0, 7, 15, 16):
        r = ope.mc_price_ex(100.0, 0.75, 0.03, 0.0, hp, ope.Payoff.European(strikes[k]), cfg, JUMP,
                            ope.McOptions(precision)).result
        assert (p17[0, k], s17[0, k]) == (r.price, r.std_error), k
    # repeated direct calls: each launch's completion counter starts from zero
    again, _, _, _ = ope.mc_surface(100.0, 0.03, 0.0, hp, JUMP, 0, [0.75], strikes[:16], cfg, precision)
    assert np.array_equal(again, p16)


def _uneven_requests(ope, n, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        v0, theta = rng.uniform(0.01, 0.16, 2)
        hp = ope.HestonParams(v0, rng.uniform(0.5, 5.0), theta, rng.uniform(0.1, 1.0), rng.uniform(-0.95, 0.0))
        kind = int(rng.integers(0, 3))
        pay = [ope.Payoff.European(100.0), ope.Payoff.AsianArithmetic(100.0), ope.Payoff.UpAndOut(100.0, 130.0)][kind]
        paths = int(np.exp(rng.uniform(np.log(1e4), np.log(3e5)))) // 2 * 2
        steps = int(rng.integers(50, 201))
        cfg = ope.McConfig(n_paths=paths, n_steps=steps, seed=1000 + i, scheme=int(rng.integers(0, 2)))
        lam = float(rng.uniform(0.0, 1.0)) if rng.random() < 0.5 else 0.0
        reqs.append(ope.McRequest(100.0, 1.0, 0.03, 0.0, hp, pay, cfg, lam, -0.1, 0.15))
    return reqs


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_equals_single_calls_bitwise(ope, precision):
    reqs = _uneven_requests(ope, 24, 7)
    res, stats, ms = ope.mc_batch(reqs, precision)
    for q, r, s in zip(reqs, res, stats):
        one = ope.mc_price_ex(q.spot, q.expiry, q.rate, q.div_yield, q.params, q.payoff, q.cfg,
                              (q.lambda_, q.mu_j, q.sigma_j), ope.McOptions(precision))
        assert (r.price, r.std_error) == (one.result.price, one.result.std_error)
        assert s.count == q.cfg.n_paths
        assert s.path_steps == q.cfg.n_paths * q.cfg.n_steps


def test_batch_order_does_not_matter(ope):
    reqs = _uneven_requests(ope, 16, 8)
    a, _, _ = ope.mc_batch(reqs, 1)
    b, _, _ = ope.mc_batch(reqs[::-1], 1)
    assert a == b[::-1]


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("n_codes", [4, 5, 6])
def test_batch_groups_on_every_stream_equal_single_calls(ope, precision, n_codes):
    # n_codes configurations (scheme x payoff) give one specialised launch
    # each, spread over the engine's streams; with 5 and 6 groups two of them
    # share a stream and the reduction runs on the last group's stream
    hp = ope.HestonParams(**BASE)
    pays = [ope.Payoff.European(100.0), ope.Payoff.AsianArithmetic(95.0), ope.Payoff.UpAndOut(100.0, 125.0)]
    reqs = []
    for i in range(n_codes):
        cfg = ope.McConfig(n_paths=20_000 + 2_000 * i, n_steps=30 + i, seed=300 + i, scheme=i // 3)
        reqs.append(ope.McRequest(100.0, 1.0, 0.03, 0.0, hp, pays[i % 3], cfg, 0.0, -0.1, 0.15))
    for order in (reqs, reqs[::-1]):
        res, stats, ms = ope.mc_batch(order, precision)
        for q, r in zip(order, res):
            one = ope.mc_price_ex(q.spot, q.expiry, q.rate, q.div_yield, q.params, q.payoff, q.cfg,
                                  (q.lambda_, q.mu_j, q.sigma_j), ope.McOptions(precision))
            assert (r.price, r.std_error) == (one.result.price, one.result.std_error)


@pytest.mark.parametrize("scheme", [0, 1])
def test_fp32_tracks_fp64_in_extreme_regimes(ope, scheme):
    # high vol-of-vol against slow mean reversion (psi far above 1.5: the QE
    # exponential branch and its point mass dominate), strong jumps, long
    # maturity, a put; FP32 and FP64 see the same draws
    cases = [
        (dict(v0=0.3, kappa=0.5, theta=0.09, sigma=2.0, rho=-0.9), 5.0, 200, (2.0, -0.2, 0.3)),
        (dict(v0=0.01, kappa=8.0, theta=0.25, sigma=1.5, rho=0.5), 0.25, 90, None),
        (dict(v0=0.04, kappa=0.2, theta=0.02, sigma=0.05, rho=-0.3), 10.0, 400, (0.1, 0.05, 0.05)),
    ]
    pays = [ope.Payoff.European(110.0, ope.OptionType.Put), ope.Payoff.AsianArithmetic(90.0),
            ope.Payoff.UpAndOut(100.0, 180.0, 2.0)]
    for hp, expiry, steps, jumps in cases:
        cfg = ope.McConfig(n_paths=200_000, n_steps=steps, seed=77, scheme=scheme)
        for pay in pays:
            f64 = ope.mc_price_ex(100.0, expiry, 0.02, 0.01, ope.HestonParams(**hp), pay, cfg, jumps,
                                  ope.McOptions(0)).result
            f32 = ope.mc_price_ex(100.0, expiry, 0.02, 0.01, ope.HestonParams(**hp), pay, cfg, jumps,
                                  ope.McOptions(1)).result
            print(f"extreme {hp} T={expiry} {pay}: fp64={f64} fp32={f32}", flush=True)
            assert math.isfinite(f32.price) and math.isfinite(f32.std_error)
            assert abs(f32.price - f64.price) <= 0.1 * f64.std_error, (hp, expiry, pay)
