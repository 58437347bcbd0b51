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


def test_seed_expansion_value_breakage_answer(ora):
    # rand_core 0.6's own test of SeedableRng::seed_from_u64 (the default the
    # reference reaches through SmallRng::seed_from_u64, mc.rs:172): the first
    # 8 bytes of the expansion of each seed below, read little endian, have
    # popcount 20..44 and differ pairwise in >= 20 bits, and for seed 0 they
    # are 5029875928683246316 (its value-breakage constant).  Forwarding to
    # Xoshiro's SplitMix64 seeding, advancing the PCG state after the output,
    # or a big-endian read would all change that word.
    st = np.zeros(4, dtype=np.uint64)
    first = []
    for seed in [0, 1, 2, 3, 4, 8, 16, 2**64 - 1]:
        ora.lib().ora_smallrng_seed(seed, st)
        first.append(int(st[0]))
    assert first[0] == 5029875928683246316
    for i, a in enumerate(first):
        assert 20 <= bin(a).count("1") <= 44
        for j, b in enumerate(first):
            if i != j:
                assert bin(a ^ b).count("1") >= 20


def test_gen_f64_is_53_bit_grid_in_unit_interval(ora):
    x = np.zeros(200_000)
    ora.lib().ora_gen_f64_n(7, x.size, x)
    assert x.min() >= 0.0 and x.max() < 1.0
    assert np.all(np.floor(x * 2.0**53) == x * 2.0**53)
    assert abs(x.mean() - 0.5) < 4 * math.sqrt(1 / 12 / x.size)
    assert stats.kstest(x, "uniform").pvalue > 1e-4


def test_gen_f64_uses_all_53_bits(ora):
    # (x >> 11) * 2^-53: below 1/2 the odd multiples of 2^-53 are exact and
    # occur half the time; a 52-bit mapping would never produce them.
    x = np.zeros(100_000)
    ora.lib().ora_gen_f64_n(9, x.size, x)
    k = (x * 2.0**53).astype(np.uint64)
    frac_odd = float(np.mean(k % 2 == 1))
    assert 0.48 < frac_odd < 0.52


def test_gen_range_respects_box_muller_bounds(ora):
    x = np.zeros(200_000)
    ora.lib().ora_gen_range_n(11, 1e-12, 1.0, x.size, x)
    assert x.min() >= 1e-12 and x.max() < 1.0
    assert stats.kstest(x, "uniform").pvalue > 1e-4


def test_gen_range_is_on_the_52_bit_grid(ora):
    # rand 0.8 UniformFloat::sample_single: value1_2 has the top 52 bits of the
    # word as its mantissa, so every output is fl(fl(k 2^-52 * scale) + low)
    # for an integer k < 2^52 with scale = 1 - 1e-12.  A 53-bit mapping
    # (j 2^-53 with j odd half the time) lands between those points.
    lo, hi = 1e-12, 1.0
    scale = hi - lo
    x = np.zeros(50_000)
    ora.lib().ora_gen_range_n(13, lo, hi, x.size, x)
    k = np.rint((x - lo) / scale * 2.0**52)
    back = (k * 2.0**-52) * scale + lo
    small = x < 0.5          # where one step of 2^-53 * scale exceeds an ulp of x
    assert small.sum() > 20_000
    assert np.array_equal(back[small], x[small])
    # and the same for one generator output worked by hand: state {1,2,3,4}
    # gives first word 41943041 (Blackman-Vigna), hence k = 41943041 >> 12
    st = np.array([1, 2, 3, 4], dtype=np.uint64)
    w = np.zeros(1, dtype=np.uint64)
    ora.lib().ora_xoshiro_next_n(st, 1, w)
    assert int(w[0]) == 41943041
    assert (41943041 >> 12) * 2.0**-52 * scale + lo == ((41943041 >> 12) * 2.0**-52) * scale + lo


class _Xoshiro256pp:
    """Xoshiro256++ (Blackman-Vigna), independent of the oracle's C code."""

    M = 2**64 - 1

    def __init__(self, s):
        self.s = [int(v) for v in s]

    @staticmethod
    def _rotl(x, k):
        return ((x << k) | (x >> (64 - k))) & _Xoshiro256pp.M

    def next(self):
        s, M = self.s, self.M
        out = (self._rotl((s[0] + s[3]) & M, 23) + s[0]) & M
        t = (s[1] << 17) & M
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = self._rotl(s[3], 45)
        return out


def _words_box_muller(w1, w2):
    # u1 from gen_range(1e-12..1.0) (52-bit mapping), u2 from gen::<f64>() (53-bit)
    lo, scale = 1e-12, 1.0 - 1e-12
    u1 = ((w1 >> 12) * 2.0**-52) * scale + lo
    u2 = (w2 >> 11) * 2.0**-53
    r = math.sqrt(-2.0 * math.log(u1))
    th = 6.283185307179586 * u2
    return r * math.cos(th), r * math.sin(th)


@pytest.mark.parametrize("lam_dt", [0.0, 1.5])
def test_draw_step_word_order(ora, lam_dt):
    # mc.rs:129-145: one draw_step consumes the Box-Muller pair (2 words) first,
    # then, with jumps, the Poisson sampler's n + 1 words (mc.rs:321-331; none
    # when lambda dt = 0), then a second Box-Muller pair (2 words) only if
    # n > 0, whose cosine output u gives jump_sum = mu_j n + sigma_j sqrt(n) u.
    # Poisson before the normals, an unconditional jump normal or a Poisson
    # loop that stops one word early would each fail this.
    n_steps, dt, mu_j, sigma_j = 400, 1.0, -0.1, 0.15
    jp = ora.jumps(lam_dt / dt, mu_j, sigma_j)
    st0 = np.zeros(4, dtype=np.uint64)
    ora.lib().ora_smallrng_seed(ora.lib().ora_splitmix64(42), st0)
    st = st0.copy()
    za, zb, js = (np.zeros(n_steps) for _ in range(3))
    cnt = np.zeros(n_steps, dtype=np.uint32)
    ora.lib().ora_draw_steps_from_state(st, jp, dt, n_steps, za, zb, js, cnt)

    g = _Xoshiro256pp(st0)
    el = math.exp(-lam_dt)
    words = 0
    for i in range(n_steps):
        a, b = _words_box_muller(g.next(), g.next())
        words += 2
        assert (a, b) == (za[i], zb[i]), i
        n, jump = 0, 0.0
        if lam_dt > 0:
            p, k = 1.0, 0
            while True:
                k += 1
                p *= (g.next() >> 11) * 2.0**-53
                words += 1
                if p <= el:
                    n = k - 1
                    break
            if n:
                u, _ = _words_box_muller(g.next(), g.next())
                words += 2
                jump = mu_j * n + sigma_j * math.sqrt(n) * u
        assert n == cnt[i], i
        assert jump == js[i], i
    assert [int(v) for v in st] == g.s
    expected = n_steps * 2 + (int(cnt.sum()) + n_steps + 2 * int((cnt > 0).sum()) if lam_dt > 0 else 0)
    assert words == expected
    if lam_dt > 0:
        assert 0.6 < (cnt > 0).mean() < 0.9   # both branches exercised


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
