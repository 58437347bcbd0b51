"""Pins the oracle's path arithmetic: the QE variance step (mc.rs:288-309),
ncdf (math.rs:16-19) and run_path (mc.rs:210-278) on given draws, against
closed forms, textbook moments and exact pathwise identities."""
import math

import numpy as np
import pytest
from scipy.special import ndtri


def cir_moments(v, kappa, theta, sigma, dt):
    """Conditional mean and variance of the CIR process after dt (textbook)."""
    e = math.exp(-kappa * dt)
    m = theta + (v - theta) * e
    s2 = v * sigma**2 * e * (1 - e) / kappa + theta * sigma**2 * (1 - e) ** 2 / (2 * kappa)
    return m, s2


@pytest.mark.parametrize("v", [0.0, 0.001, 0.04, 0.2, 0.5, 1.5])
def test_qe_step_matches_cir_moments(ora, v):
    kappa, theta, sigma, dt = 5.0, 0.04, 1.2, 1.0 / 8
    n = 2_000_000
    u = (np.arange(n) + 0.5) / n
    z = ndtri(u)
    out = ora.qe_variance_step(np.full(n, v), z, kappa, theta, sigma, dt)
    m, s2 = cir_moments(v, kappa, theta, sigma, dt)
    psi = s2 / m**2
    assert np.all(out >= 0.0)
    assert abs(out.mean() - m) < 2e-4 * m
    assert abs(out.var() - s2) < 2e-3 * s2
    if psi > 1.5:
        # exponential branch: point mass p = (psi-1)/(psi+1) at zero
        p = (psi - 1) / (psi + 1)
        assert abs((out == 0).mean() - p) < 2.0 / n + 1e-6
    else:
        # quadratic branch: v' = a (b + z)^2 is increasing in z around z > -b
        assert (out == 0).mean() < 1e-6


def test_qe_step_both_branches_exercised(ora):
    kappa, theta, sigma, dt = 5.0, 0.04, 1.2, 1.0 / 8
    for v, quad in [(0.04, False), (0.5, True)]:
        m, s2 = cir_moments(v, kappa, theta, sigma, dt)
        assert (s2 / m**2 <= 1.5) == quad


def test_ncdf_reference_values(ora):
    # math.rs:149-168
    assert abs(ora.lib().ora_ncdf(0.0) - 0.5) < 1e-15
    assert abs(ora.lib().ora_ncdf(1.64485363) - 0.95) < 1e-6
    assert ora.lib().ora_ncdf(-10.0) < 1e-23
    for x, want in [(-4.0, 3.167124183311998e-5), (-5.0, 2.866515718791939e-7), (-6.0, 9.865876449133282e-10)]:
        got = ora.lib().ora_ncdf(x)
        assert abs(got - want) / want < 5e-10


# ----------------------------------------------------------- run_path

S0, T, R, Q = 100.0, 1.0, 0.05, 0.0


def draws(n_pairs, n_steps, seed=0, jumps=False):
    rng = np.random.default_rng(seed)
    za = rng.standard_normal((n_steps, n_pairs))
    zb = rng.standard_normal((n_steps, n_pairs))
    js = np.where(rng.random((n_steps, n_pairs)) < 0.05, rng.normal(-0.1, 0.15, (n_steps, n_pairs)), 0.0)
    return za, zb, (js if jumps else np.zeros_like(za))


def test_zero_volvol_euler_is_exact_gbm(ora):
    theta = 0.04
    hp = ora.heston(theta, 1.5, theta, 0.0, -0.7)
    n_steps = 17
    za, zb, js = draws(64, n_steps, 1)
    pay, st = ora.run_paths(S0, T, R, Q, hp, ora.payoff(), n_steps, ora.EULER, True, za, zb, js)
    dt = T / n_steps
    ln_t = math.log(S0) + n_steps * (R - Q - theta / 2) * dt + math.sqrt(theta) * math.sqrt(dt) * za.sum(0)
    assert np.all(st[:, 0, 1] == theta)
    np.testing.assert_allclose(st[:, 0, 0], ln_t, rtol=0, atol=1e-12)
    ln_b = math.log(S0) + n_steps * (R - Q - theta / 2) * dt - math.sqrt(theta) * math.sqrt(dt) * za.sum(0)
    np.testing.assert_allclose(st[:, 1, 0], ln_b, rtol=0, atol=1e-12)
    np.testing.assert_allclose(pay[:, 0], math.exp(-R * T) * np.maximum(np.exp(ln_t) - 100.0, 0), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("scheme", [0, 1])
def test_asian_one_step_equals_european(ora, scheme):
    hp = ora.heston(0.04, 2.0, 0.04, 0.3, -0.7)
    za, zb, js = draws(200, 1, 2, jumps=True)
    jp = ora.jumps(0.5, -0.1, 0.15)
    eu, _ = ora.run_paths(S0, T, R, Q, hp, ora.payoff(ora.EUROPEAN, 95.0), 1, scheme, True, za, zb, js, jp)
    asi, _ = ora.run_paths(S0, T, R, Q, hp, ora.payoff(ora.ASIAN, 95.0), 1, scheme, True, za, zb, js, jp)
    assert np.array_equal(eu, asi)


@pytest.mark.parametrize("scheme", [0, 1])
def test_barrier_limits(ora, scheme):
    hp = ora.heston(0.04, 2.0, 0.04, 0.3, -0.7)
    za, zb, js = draws(300, 12, 3)
    eu, _ = ora.run_paths(S0, T, R, Q, hp, ora.payoff(ora.EUROPEAN, 100.0), 12, scheme, True, za, zb, js)
    far, _ = ora.run_paths(S0, T, R, Q, hp, ora.payoff(ora.UP_AND_OUT, 100.0, barrier=1e300), 12, scheme, True,
                           za, zb, js)
    assert np.array_equal(eu, far)
    # barrier at or below spot: knocked out at t=0 (running_max starts at spot),
    # rebate discounted from T (mc.rs:218, 276-277)
    ko, st = ora.run_paths(S0, T, R, Q, hp, ora.payoff(ora.UP_AND_OUT, 100.0, barrier=100.0, rebate=2.5), 12,
                           scheme, True, za, zb, js)
    assert np.all(ko == math.exp(-R * T) * 2.5)
    assert np.all(st[:, :, 3] >= S0)


@pytest.mark.parametrize("scheme", [0, 1])
def test_antithetic_sibling_is_negated_draws(ora, scheme):
    hp = ora.heston(0.04, 2.0, 0.04, 0.3, -0.7)
    za, zb, js = draws(100, 9, 4, jumps=True)
    jp = ora.jumps(0.5, -0.1, 0.15)
    p1, s1 = ora.run_paths(S0, T, R, Q, hp, ora.payoff(ora.ASIAN, 100.0), 9, scheme, True, za, zb, js, jp)
    p2, s2 = ora.run_paths(S0, T, R, Q, hp, ora.payoff(ora.ASIAN, 100.0), 9, scheme, True, -za, -zb, js, jp)
    assert np.array_equal(p1[:, 1], p2[:, 0])
    assert np.array_equal(s1[:, 1], s2[:, 0])


@pytest.mark.parametrize("scheme", [0, 1])
def test_jumps_shift_log_price_only(ora, scheme):
    hp = ora.heston(0.04, 2.0, 0.04, 0.3, -0.7)
    za, zb, js = draws(100, 15, 5, jumps=True)
    jp = ora.jumps(0.5, -0.1, 0.15)
    nj = ora.jumps(0.5, -0.1, 0.15)
    _, s_j = ora.run_paths(S0, T, R, Q, hp, ora.payoff(), 15, scheme, True, za, zb, js, jp)
    _, s_0 = ora.run_paths(S0, T, R, Q, hp, ora.payoff(), 15, scheme, True, za, zb, np.zeros_like(js), nj)
    assert np.array_equal(s_j[:, :, 1], s_0[:, :, 1])  # variance path unchanged
    np.testing.assert_allclose(s_j[:, 0, 0] - s_0[:, 0, 0], js.sum(0), atol=1e-12)


@pytest.mark.parametrize("scheme", [0, 1])
def test_call_minus_put_is_discounted_forward_pathwise(ora, scheme):
    hp = ora.heston(0.04, 2.0, 0.04, 0.3, -0.7)
    za, zb, js = draws(150, 10, 6)
    c, st = ora.run_paths(S0, T, R, Q, hp, ora.payoff(ora.EUROPEAN, 103.0, ora.CALL), 10, scheme, True, za, zb, js)
    p, _ = ora.run_paths(S0, T, R, Q, hp, ora.payoff(ora.EUROPEAN, 103.0, ora.PUT), 10, scheme, True, za, zb, js)
    np.testing.assert_allclose(c - p, math.exp(-R * T) * (np.exp(st[:, :, 0]) - 103.0), rtol=0, atol=1e-10)


def test_gen_draws_follow_chunk_seeding(ora):
    # chunk c draws from SmallRng::seed_from_u64(splitmix64(seed ^ c)) (mc.rs:172):
    # with one thread, pair p's draws are the p-th block of one stream; with
    # P threads and P pairs each pair starts a fresh stream.
    za1, zb1, _ = ora.gen_draws(42, 4, 3, 1.0, n_threads=1)
    za4, zb4, _ = ora.gen_draws(42, 4, 3, 1.0, n_threads=4)
    assert np.array_equal(za1[:, 0], za4[:, 0])
    assert not np.array_equal(za1[:, 1], za4[:, 1])
    z = np.zeros(6)
    L = ora.lib()
    L.ora_box_muller_n(L.ora_splitmix64(42 ^ 2), 3, z)
    assert np.array_equal(za4[:, 2], z[0::2])
    assert np.array_equal(zb4[:, 2], z[1::2])
