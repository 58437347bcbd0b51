"""Pins the oracle's random-number layer.

Reference semantics: mc.rs:172 (SmallRng::seed_from_u64(splitmix64(seed ^
chunk))), mc.rs:312-318 (Box-Muller), mc.rs:321-331 (Knuth Poisson),
mc.rs:337-343 (splitmix64).  The generators are pinned by published
known-answer vectors; the samplers by their distributions.
"""
import math

import numpy as np
import pytest
from scipy import stats


def test_splitmix64_known_answers(ora):
    # SplitMix64 (Vigna) from state 0: the reference's splitmix64(x) adds the
    # golden gamma then mixes, i.e. it is the generator's output for state x.
    L = ora.lib()
    assert L.ora_splitmix64(0) == 0xE220A8397B1DCDAF
    assert L.ora_splitmix64(0x9E3779B97F4A7C15) == 0x6E789E6AA1B965F4
    assert L.ora_splitmix64((2 * 0x9E3779B97F4A7C15) % 2**64) == 0x06C45D188009454F


def test_xoshiro256pp_known_answers(ora):
    # Blackman-Vigna xoshiro256++ from state {1, 2, 3, 4}.
    st = np.array([1, 2, 3, 4], dtype=np.uint64)
    out = np.zeros(10, dtype=np.uint64)
    ora.lib().ora_xoshiro_next_n(st, 10, out)
    expected = [41943041, 58720359, 3588806011781223, 3591011842654386, 9228616714210784205,
                9973669472204895162, 14011001112246962877, 12406186145184390807,
                15849039046786891736, 10450023813501588000]
    assert [int(x) for x in out] == expected


def test_pcg32_output_function_known_answers(ora):
    # Canonical pcg32 demo (initstate 42, initseq 54) pins the XSH-RR output
    # permutation that rand_core's seed_from_u64 expansion uses.
    out = np.zeros(6, dtype=np.uint32)
    ora.lib().ora_pcg32_canonical(42, 54, 6, out)
    assert [int(x) for x in out] == [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]


def test_seed_expansion_is_deterministic_and_spread(ora):
    st = np.zeros(4, dtype=np.uint64)
    seen = set()
    for seed in [0, 1, 2, 3, 4, 8, 16, 2**64 - 1]:
        ora.lib().ora_smallrng_seed(seed, st)
        words = tuple(int(x) for x in st)
        assert words not in seen
        seen.add(words)
        # every 64-bit word of a PCG32 expansion has a plausible popcount
        for w in words:
            assert 12 <= bin(w).count("1") <= 52


def test_gen_f64_is_53_bit_grid_in_unit_interval(ora):
    x = np.zeros(200_000)
    ora.lib().ora_gen_f64_n(7, x.size, x)
    assert x.min() >= 0.0 and x.max() < 1.0
    assert np.all(np.floor(x * 2.0**53) == x * 2.0**53)
    assert abs(x.mean() - 0.5) < 4 * math.sqrt(1 / 12 / x.size)
    assert stats.kstest(x, "uniform").pvalue > 1e-4


def test_gen_range_respects_box_muller_bounds(ora):
    x = np.zeros(200_000)
    ora.lib().ora_gen_range_n(11, 1e-12, 1.0, x.size, x)
    assert x.min() >= 1e-12 and x.max() < 1.0
    assert stats.kstest(x, "uniform").pvalue > 1e-4


def test_box_muller_normals(ora):
    n = 500_000
    z = np.zeros(2 * n)
    ora.lib().ora_box_muller_n(3, n, z)
    a, b = z[0::2], z[1::2]
    # |z| <= sqrt(-2 ln 1e-12) because u1 >= 1e-12 (mc.rs:313)
    assert np.abs(z).max() <= math.sqrt(-2 * math.log(1e-12)) + 1e-12
    for v in (a, b):
        assert abs(v.mean()) < 5 / math.sqrt(n)
        assert abs(v.var() - 1) < 5 * math.sqrt(2 / n)
        assert abs(stats.kurtosis(v)) < 0.05
        assert stats.kstest(v, "norm").pvalue > 1e-4
    assert abs(np.corrcoef(a, b)[0, 1]) < 5 / math.sqrt(n)


def test_poisson_sampler_matches_mean(ora):
    # mc.rs:457-465: SmallRng::seed_from_u64(1), mean 0.3, 200k draws, +/- 0.01
    n = 200_000
    out = np.zeros(n, dtype=np.uint32)
    ora.lib().ora_sample_poisson_n(1, 0.3, n, out)
    assert abs(out.mean() - 0.3) < 0.01


@pytest.mark.parametrize("mean", [0.3, 2.5])
def test_poisson_sampler_pmf(ora, mean):
    n = 300_000
    out = np.zeros(n, dtype=np.uint32)
    ora.lib().ora_sample_poisson_n(5, mean, n, out)
    kmax = 12
    obs = np.bincount(np.minimum(out, kmax), minlength=kmax + 1)
    p = stats.poisson.pmf(np.arange(kmax), mean)
    p = np.append(p, 1 - p.sum())
    keep = p * n > 5
    chi2 = ((obs[keep] - n * p[keep]) ** 2 / (n * p[keep])).sum()
    assert stats.chi2.sf(chi2, keep.sum() - 1) > 1e-4


def test_poisson_zero_mean_consumes_nothing(ora):
    out = np.ones(10, dtype=np.uint32)
    ora.lib().ora_sample_poisson_n(1, 0.0, 10, out)
    assert np.all(out == 0)


# --- the GPU engine's Philox stream specification ------------------------

def test_philox4x32_10_known_answers(ora):
    # Random123 known-answer vectors for philox4x32 with 10 rounds.
    cases = [
        ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
        ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
        ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
         (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
    ]
    for ctr, key, want in cases:
        got = ora.philox4x32_10(*[[c] for c in ctr], key[0], key[1])
        assert tuple(int(g[0]) for g in got) == want


def test_philox_draws_are_normal_and_poisson(ora):
    za, zb, js, nn = ora.philox_draws(1234, np.arange(4096), 8, dt=1.0 / 8, lam=2.0, mu_j=-0.1, sigma_j=0.15)
    z = np.concatenate([za.ravel(), zb.ravel()])
    assert stats.kstest(z, "norm").pvalue > 1e-4
    assert abs(np.corrcoef(za.ravel(), zb.ravel())[0, 1]) < 0.03
    mu = 2.0 / 8
    assert abs(nn.mean() - mu) < 5 * math.sqrt(mu / nn.size)
    # jump sizes given n are N(n mu_j, n sigma_j^2)
    one = js[nn == 1]
    assert abs(one.mean() + 0.1) < 5 * 0.15 / math.sqrt(one.size)
    assert np.all(js[nn == 0] == 0.0)


def test_poisson_threshold_matches_cdf(ora):
    for mu in [1e-9, 0.002, 0.5, 3.0, 40.0]:
        thr = ora.poisson_threshold0(mu)
        f0 = math.exp(-mu)
        assert (thr + 0.5) * 2.0**-32 >= f0 or thr >= 2**32 - 1
        assert thr == 0 or (thr - 1 + 0.5) * 2.0**-32 < f0
