"""Replay oracle: the GPU path kernels fed with the reference's own draws
(the oracle's restatement of mc.rs's SmallRng stream) must reproduce the
oracle's per-path results.

Tolerance contract (FP64):
  * path state (ln S_T, v_T, running sum, running max) to 1e-12 relative;
  * discounted payoffs to 1e-12 * max(|payoff|, S0) absolute;
  * knock-out decisions identical except where |running_max - B| <= 1e-12 B;
  * price and std_error to 1e-12 relative.
FP32-path mode is held to the looser bound its precision allows.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HP = dict(v0=0.04, kappa=1.5, theta=0.04, sigma=0.5, rho=-0.7)
S0, T, R, Q = 100.0, 1.0, 0.03, 0.01
JUMP = (0.5, -0.1, 0.15)


def _oracle_and_gpu(ope, ora, scheme, lam, kind, anti, opt, n_pairs, n_steps, seed=42, precision=0, strike=100.0,
                    barrier=130.0, rebate=0.0, jump_params=JUMP, hp=HP, expiry=T, rate=R, q=Q):
    jp = ora.jumps(lam, *jump_params[1:]) if lam is not None else ora.jumps()
    za, zb, js = ora.gen_draws(seed, n_pairs, n_steps, expiry, jp, n_threads=4)
    opay = ora.payoff(kind, strike, opt, barrier, rebate)
    pay_ref, st_ref = ora.run_paths(S0, expiry, rate, q, ora.heston(**hp), opay, n_steps, scheme, anti, za, zb, js,
                                    jp)
    params = ope.HestonParams(**hp)
    payoff = ope.Payoff(kind, strike, opt, barrier, rebate)
    cfg = ope.McConfig(n_paths=n_pairs * (2 if anti else 1), n_steps=n_steps, seed=seed, antithetic=anti,
                       scheme=scheme)
    jumps = (lam, jump_params[1], jump_params[2]) if lam is not None else None
    print(f"replay case scheme={scheme} lam={lam} kind={kind} anti={anti} opt={opt} pairs={n_pairs} "
          f"steps={n_steps} precision={precision} hp={hp} T={expiry} r={rate} q={q}", flush=True)
    run = ope.mc_price_ex(S0, expiry, rate, q, params, payoff, cfg, jumps,
                          ope.McOptions(precision, replay=(za, zb, js if lam else None), want_paths=True))
    return pay_ref, st_ref, run


def _ref_price(pay_ref, anti):
    p = pay_ref if anti else pay_ref[:, :1]
    n = p.size
    mean = p.sum() / n
    var = max((p * p).sum() / n - mean * mean, 0.0)
    return mean, math.sqrt(var / n)


CASES = [(s, lam, k, anti, opt)
         for s in (0, 1)
         for lam in (None, 0.5)
         for k in (0, 1, 2)
         for anti, opt in ((True, 0), (False, 1))]


@pytest.mark.parametrize("scheme,lam,kind,anti,opt", CASES)
def test_replay_fp64_matches_oracle_per_path(ope, ora, scheme, lam, kind, anti, opt):
    n_pairs, n_steps = 700, 37   # ragged last tile, odd step count
    pay_ref, st_ref, run = _oracle_and_gpu(ope, ora, scheme, lam, kind, anti, opt, n_pairs, n_steps)
    sibs = 2 if anti else 1
    st = run.path_states[:, :sibs]
    sr = st_ref[:, :sibs]
    # ln S_T and v_T, always tracked
    for c in (0, 1):
        np.testing.assert_allclose(st[..., c], sr[..., c], rtol=1e-12, atol=1e-300)
    if kind == 1:
        np.testing.assert_allclose(st[..., 2], sr[..., 2], rtol=1e-12)
    if kind == 2:
        np.testing.assert_allclose(st[..., 3], sr[..., 3], rtol=1e-12)
    pay = run.path_payoffs[:, :sibs]
    pr = pay_ref[:, :sibs]
    tol = 1e-12 * np.maximum(np.abs(pr), S0)
    bad = np.abs(pay - pr) > tol
    if kind == 2:
        near = np.abs(sr[..., 3] - 130.0) <= 1e-12 * 130.0
        bad &= ~near
    assert not bad.any(), (np.argwhere(bad)[:5], pay[bad][:5], pr[bad][:5])
    price, se = _ref_price(pay_ref, anti)
    assert abs(run.result.price - price) <= 1e-12 * abs(price) + 1e-15
    assert abs(run.result.std_error - se) <= 1e-12 * se + 1e-15
    assert run.stats.count == n_pairs * sibs
    assert run.stats.path_steps == n_pairs * sibs * n_steps


def test_replay_fp64_state_is_bitwise_for_euler(ope, ora):
    # Euler's state update uses only +, *, sqrt and max, all correctly rounded:
    # with the same unfused operation order the state agrees bit for bit.
    _, st_ref, run = _oracle_and_gpu(ope, ora, 0, 0.5, 0, True, 0, 513, 64)
    assert np.array_equal(run.path_states[..., 0], st_ref[..., 0])
    assert np.array_equal(run.path_states[..., 1], st_ref[..., 1])


@pytest.mark.parametrize("n_pairs,n_steps", [(1, 1), (1, 5), (256, 1), (257, 2)])
def test_replay_edge_shapes(ope, ora, n_pairs, n_steps):
    for scheme in (0, 1):
        pay_ref, st_ref, run = _oracle_and_gpu(ope, ora, scheme, 0.5, 2, True, 0, n_pairs, n_steps, barrier=101.0)
        np.testing.assert_allclose(run.path_states[..., :2], st_ref[..., :2], rtol=1e-12)
        np.testing.assert_allclose(run.path_payoffs, pay_ref, rtol=0, atol=1e-12 * S0)


def test_replay_feller_violating_qe_exercises_both_branches(ope, ora):
    # kappa=5, theta=.04, sigma=1.2, 8 steps (mc.rs:373): psi > 1.5 is common
    hp = dict(v0=0.04, kappa=5.0, theta=0.04, sigma=1.2, rho=-0.7)
    pay_ref, st_ref, run = _oracle_and_gpu(ope, ora, 1, None, 0, True, 0, 2000, 8, hp=hp)
    np.testing.assert_allclose(run.path_states[..., :2], st_ref[..., :2], rtol=1e-12, atol=1e-300)
    assert (st_ref[..., 1] == 0).mean() > 0.01   # point mass of the exponential branch reached
    np.testing.assert_allclose(run.path_payoffs, pay_ref, rtol=0, atol=1e-12 * S0)


@pytest.mark.parametrize("scheme,lam,kind", [(0, None, 0), (0, 0.5, 1), (1, None, 2), (1, 0.5, 0)])
def test_replay_fp32_path_close_to_oracle(ope, ora, scheme, lam, kind):
    pay_ref, st_ref, run = _oracle_and_gpu(ope, ora, scheme, lam, kind, True, 0, 2000, 50, precision=1)
    st = run.path_states
    np.testing.assert_allclose(st[..., 0], st_ref[..., 0], rtol=0, atol=2e-5)
    np.testing.assert_allclose(st[..., 1], st_ref[..., 1], rtol=0, atol=2e-5)
    d = np.abs(run.path_payoffs - pay_ref)
    frac_close = (d <= 5e-3).mean()
    assert frac_close > 0.998
    price, se = _ref_price(pay_ref, True)
    assert abs(run.result.price - price) < 0.02 * se


@pytest.mark.parametrize("scheme,lam,kind", [(1, None, 0), (1, None, 1), (1, 0.5, 2), (0, None, 1), (0, 0.5, 2)])
def test_replay_fp32_path_without_sibling(ope, ora, scheme, lam, kind):
    # antithetic = false: one path per draw column; the FP32 QE loop runs
    # without the sibling's code (a separately compiled case)
    pay_ref, st_ref, run = _oracle_and_gpu(ope, ora, scheme, lam, kind, False, 1, 2000, 50, precision=1)
    st = run.path_states[:, :1]
    np.testing.assert_allclose(st[..., 0], st_ref[:, :1, 0], rtol=0, atol=2e-5)
    np.testing.assert_allclose(st[..., 1], st_ref[:, :1, 1], rtol=0, atol=2e-5)
    assert np.isnan(run.path_states[:, 1]).all() and np.isnan(run.path_payoffs[:, 1]).all()
    d = np.abs(run.path_payoffs[:, 0] - pay_ref[:, 0])
    assert (d <= 5e-3).mean() > 0.998
    price, se = _ref_price(pay_ref, False)
    assert abs(run.result.price - price) < 0.02 * se


@pytest.mark.parametrize("scheme,lam,kind,n_steps", [(0, None, 0, 37), (0, 0.5, 0, 1), (0, 0.5, 1, 37),
                                                     (0, None, 2, 9), (1, 0.5, 0, 37)])
@pytest.mark.parametrize("precision", [0, 1])
def test_replay_negative_v0(ope, ora, scheme, lam, kind, n_steps, precision):
    # v0 < 0: the first step sees v+ = 0 (mc.rs:245): no diffusion, and Euler's
    # v' = max(v0 + kappa theta dt, 0) (mc.rs:254); QE starts from v+ = 0
    hp = dict(HP, v0=-0.02)
    pay_ref, st_ref, run = _oracle_and_gpu(ope, ora, scheme, lam, kind, True, 0, 600, n_steps, hp=hp,
                                           precision=precision)
    st = run.path_states
    if precision == 0:
        np.testing.assert_allclose(st[..., :2], st_ref[..., :2], rtol=1e-12, atol=1e-300)
        np.testing.assert_allclose(run.path_payoffs, pay_ref, rtol=0, atol=1e-12 * S0)
    else:
        np.testing.assert_allclose(st[..., 0], st_ref[..., 0], rtol=0, atol=2e-5)
        np.testing.assert_allclose(st[..., 1], st_ref[..., 1], rtol=0, atol=2e-5)
        assert (np.abs(run.path_payoffs - pay_ref) <= 5e-3).mean() > 0.998


EDGE_PARAMS = [
    # (id, Heston overrides, expiry, rate, dividend yield, scheme, lambda, payoff kind, steps)
    ("rho_minus_1_euler", dict(rho=-1.0), 1.0, 0.03, 0.01, 0, None, 0, 24),
    ("rho_minus_1_qe_bates_asian", dict(rho=-1.0), 1.0, 0.03, 0.01, 1, 0.5, 1, 24),
    ("rho_plus_1_qe_uao", dict(rho=1.0), 1.0, 0.03, 0.01, 1, None, 2, 24),
    ("rho_plus_1_euler_bates_uao", dict(rho=1.0), 1.0, 0.03, 0.01, 0, 0.5, 2, 24),
    ("negative_rate_dividend_above_rate", {}, 2.0, -0.01, 0.05, 0, 0.5, 2, 30),
    ("short_expiry_qe_bates", {}, 1e-3, 0.03, 0.0, 1, 0.5, 0, 3),
    ("fast_mean_reversion_qe", dict(kappa=50.0), 1.0, 0.03, 0.0, 1, None, 1, 4),
]


@pytest.mark.parametrize("case", EDGE_PARAMS, ids=[c[0] for c in EDGE_PARAMS])
def test_replay_fp64_edge_parameters(ope, ora, case):
    # |rho| = 1 (no independent diffusion: sqrt(1 - rho^2) = 0, QE's K3 = K4
    # = 0), a negative rate with q above r, a 1e-3-year expiry in 3 steps,
    # kappa dt = 12.5 (QE's e^{-kappa dt} ~ 4e-6): per-path parity as above
    _, over, expiry, rate, q, scheme, lam, kind, n_steps = case
    hp = dict(HP, **over)
    pay_ref, st_ref, run = _oracle_and_gpu(ope, ora, scheme, lam, kind, True, 0, 700, n_steps, hp=hp,
                                           expiry=expiry, rate=rate, q=q)
    st = run.path_states
    for c in (0, 1):
        np.testing.assert_allclose(st[..., c], st_ref[..., c], rtol=1e-12, atol=1e-300)
    if kind == 1:
        np.testing.assert_allclose(st[..., 2], st_ref[..., 2], rtol=1e-12)
    tol = 1e-12 * np.maximum(np.abs(pay_ref), S0)
    bad = np.abs(run.path_payoffs - pay_ref) > tol
    if kind == 2:
        np.testing.assert_allclose(st[..., 3], st_ref[..., 3], rtol=1e-12)
        bad &= ~(np.abs(st_ref[..., 3] - 130.0) <= 1e-12 * 130.0)
    assert not bad.any(), (np.argwhere(bad)[:5], run.path_payoffs[bad][:5], pay_ref[bad][:5])
    price, se = _ref_price(pay_ref, True)
    assert abs(run.result.price - price) <= 1e-12 * abs(price) + 1e-15
    assert abs(run.result.std_error - se) <= 1e-12 * se + 1e-15
