"""The FP64 Philox-mode path's elementary functions (fp64_math.h) against
exact values (mpmath at 60 digits, rounded once): within 2 ulp over the
ranges the path uses and beyond them (the normal tail within 8 ulp).  The replay kernels do not use these
(they call the device library in the reference's order)."""
import math

import mpmath
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

mpmath.mp.dps = 60


def ulp_err(got, exact):
    got = np.asarray(got, dtype=np.float64)
    ex = np.asarray([float(v) for v in exact])
    return np.abs(got - ex) / np.spacing(np.abs(ex))


def test_log(ope):
    rng = np.random.default_rng(1)
    # the Box-Muller uniform (w + 1/2) 2^-32 over all w, values next to 1,
    # and the QE exponential branch's arguments (t / (1 - ncdf), up to ~1e12)
    w = np.concatenate([rng.integers(0, 2**32, 4000), np.arange(0, 64), 2**32 - 1 - np.arange(0, 64)])
    x = np.concatenate([(w.astype(np.float64) + 0.5) * 2.0**-32,
                        1.0 + rng.uniform(-1e-9, 1e-9, 500), 10.0 ** rng.uniform(-300, 300, 2000),
                        np.array([0.70710678118654746, 0.70710678118654757, 1.0, 2.0, 0.5])])
    got = ope.debug_fp64_math(0, x)
    err = ulp_err(got, [mpmath.log(mpmath.mpf(v)) for v in x])
    print(f"log: max {err.max():.2f} ulp", flush=True)
    assert np.all(got[x == 1.0] == 0.0)
    assert err.max() <= 2.0


def test_exp(ope):
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.uniform(-1.0, 1.0, 3000), rng.uniform(-700.0, 700.0, 3000),
                        rng.uniform(3.0, 6.0, 2000),           # ln S of the Asian sums
                        np.arange(-20, 21) * math.log(2.0) / 2, np.array([0.0])])
    got = ope.debug_fp64_math(1, x)
    err = ulp_err(got, [mpmath.exp(mpmath.mpf(v)) for v in x])
    print(f"exp: max {err.max():.2f} ulp", flush=True)
    assert got[-1] == 1.0
    assert err.max() <= 2.0


def test_sincos_of_the_23_bit_angle(ope):
    rng = np.random.default_rng(3)
    q = 1 << 21
    # random angles, every quarter-turn boundary and its neighbours, the eighth turns
    k = np.unique(np.concatenate([rng.integers(0, 2**23, 6000), np.arange(0, 2**23, q // 8),
                                  *[np.arange(max(j * q - 3, 0), min(j * q + 4, 2**23)) for j in range(5)]]))
    x = k.astype(np.float64)
    s = ope.debug_fp64_math(2, x)
    c = ope.debug_fp64_math(3, x)
    ang = [2 * mpmath.pi * int(v) / 2**23 for v in k]
    es, ec = [mpmath.sin(a) for a in ang], [mpmath.cos(a) for a in ang]
    # absolute error in units of 2^-53 (an ulp of values in [1/2, 1)); near the
    # zeros of sine and cosine, relative error in ulp of the value
    abs_s = np.abs(s - np.array([float(v) for v in es])) * 2.0**53
    abs_c = np.abs(c - np.array([float(v) for v in ec])) * 2.0**53
    print(f"sin/cos: max {abs_s.max():.2f} / {abs_c.max():.2f} units of 2^-53", flush=True)
    assert abs_s.max() <= 2.0 and abs_c.max() <= 2.0
    # exact zeros: sine at the half turns, cosine at the odd quarter turns
    zs, zc = k % (1 << 22) == 0, k % (1 << 22) == (1 << 21)
    assert np.all(s[zs] == 0.0) and np.all(c[zc] == 0.0)
    small = (np.abs(s) < 0.5) & ~zs
    assert ulp_err(s[small], [v for v, m in zip(es, small) if m]).max() <= 2.0
    small = (np.abs(c) < 0.5) & ~zc
    assert ulp_err(c[small], [v for v, m in zip(ec, small) if m]).max() <= 2.0
    assert np.all(c[k == 0] == 1.0)


def test_reciprocal(ope):
    rng = np.random.default_rng(4)
    x = np.concatenate([10.0 ** rng.uniform(-300, 300, 3000), -(10.0 ** rng.uniform(-10, 10, 500)),
                        np.array([1.0, 3.0, 0.1])])
    got = ope.debug_fp64_math(4, x)
    err = ulp_err(got, [1 / mpmath.mpf(v) for v in x])
    print(f"rcp: max {err.max():.2f} ulp", flush=True)
    assert err.max() <= 1.0


def test_reciprocal_square_root(ope):
    rng = np.random.default_rng(5)
    x = np.concatenate([10.0 ** rng.uniform(-300, 300, 3000), rng.uniform(0.2, 3.0, 2000),
                        np.array([1.0, 4.0, 0.25, 2.0])])
    got = ope.debug_fp64_math(5, x)
    err = ulp_err(got, [1 / mpmath.sqrt(mpmath.mpf(v)) for v in x])
    print(f"rsqrt: max {err.max():.2f} ulp", flush=True)
    assert err.max() <= 1.0


def test_square_root(ope):
    rng = np.random.default_rng(6)
    # QE's m w = sqrt(m^2 - s^2/2) and variance terms, Euler's sqrt(v),
    # the Box-Muller radius, the jump count, and the full normal range
    x = np.concatenate([10.0 ** rng.uniform(-300, 300, 3000), rng.uniform(0.0, 1e-3, 2000),
                        rng.uniform(0.0, 46.0, 2000), np.arange(1.0, 65.0),
                        np.array([0.0, 1.0, 4.0, 0.25, 2.0, 2.2250738585072014e-308])])
    got = ope.debug_fp64_math(6, x)
    err = ulp_err(got[x > 0], [mpmath.sqrt(mpmath.mpf(v)) for v in x[x > 0]])
    print(f"sqrt: max {err.max():.2f} ulp, {np.mean(err == 0):.4f} correctly rounded", flush=True)
    assert np.all(got[x == 0.0] == 0.0)
    sq = np.array([1.0, 4.0, 0.25, 4096.0])
    assert np.all(ope.debug_fp64_math(6, sq) == np.sqrt(sq))
    assert err.max() <= 1.0


def test_normal_tail(ope):
    rng = np.random.default_rng(7)
    # every Box-Muller normal lies in [-6.77, 6.77]: the QE exponential branch's
    # ncdf(-|z|), the small tail probability that must keep its relative accuracy
    z = np.concatenate([rng.uniform(-6.8, 6.8, 4000), np.linspace(0.0, 6.8, 681), np.array([0.0, -6.8, 6.8])])
    got = ope.debug_fp64_math(7, z)
    exact = [mpmath.erfc(abs(mpmath.mpf(v)) / mpmath.sqrt(2)) / 2 for v in z]
    err = ulp_err(got, exact)
    # about 3e-16 absolute at the worst points (values just below 1/2, where
    # the roundings of t and of e^{-z^2/2} in [1/2, 1) count double)
    print(f"ncdf(-|z|): max {err.max():.2f} ulp", flush=True)
    assert err.max() <= 8.0
    assert np.array_equal(got, ope.debug_fp64_math(7, -z))  # a function of |z|
    with pytest.raises(ope.OpeError):
        ope.debug_fp64_math(7, np.array([6.9]))


def test_angle_must_be_an_integer(ope):
    with pytest.raises(ope.OpeError):
        ope.debug_fp64_math(2, np.array([0.5]))
