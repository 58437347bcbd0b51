"""ORACLE -- TEST INFRASTRUCTURE ONLY.

Python access to the CPU oracle:

* ``liboracle.so`` (``mc_ref.c``, ``analytic.c``): the C restatement of the
  reference engine ``src/mc.rs`` (draw stream, ``run_path``, ``mc_price``) and
  of its analytic Heston/Bates/BSM pricers (``src/heston.rs``,
  ``src/bates.rs``, ``src/bsm.rs``).  File:line citations are in those
  sources.
* ``philox_*``: a numpy restatement of the GPU engine's own counter-based
  Philox4x32-10 draw specification (the reference has no counterpart: its
  draws come from rand's SmallRng).  It is written from the Random123 paper
  (Salmon et al., SC'11) and pinned by that paper's known-answer vectors.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this package, and only as the checker or as the timed CPU
baseline.  The product library never calls it.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

EUROPEAN, ASIAN, UP_AND_OUT = 0, 1, 2
CALL, PUT = 0, 1
EULER, QE = 0, 1


class Heston(C.Structure):
    _fields_ = [("v0", C.c_double), ("kappa", C.c_double), ("theta", C.c_double),
                ("sigma", C.c_double), ("rho", C.c_double)]


class PayoffS(C.Structure):
    _fields_ = [("kind", C.c_int32), ("opt_type", C.c_int32), ("strike", C.c_double),
                ("barrier", C.c_double), ("rebate", C.c_double)]


class Config(C.Structure):
    _fields_ = [("n_paths", C.c_uint64), ("n_steps", C.c_uint64), ("seed", C.c_uint64),
                ("antithetic", C.c_int32), ("scheme", C.c_int32)]


class Jumps(C.Structure):
    _fields_ = [("has_jumps", C.c_int32), ("lambda_", C.c_double), ("mu_j", C.c_double),
                ("sigma_j", C.c_double)]


_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")


def lib():
    """Load liboracle.so (built by ``make -C oracle`` / ``__graft_entry__.build()``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
    L = C.CDLL(LIB_PATH)
    L.ora_build_flags.restype = C.c_char_p
    L.ora_build_flags.argtypes = []
    L.ora_splitmix64.restype = C.c_uint64
    L.ora_splitmix64.argtypes = [C.c_uint64]
    L.ora_smallrng_seed.argtypes = [C.c_uint64, _u64p]
    L.ora_xoshiro_next_n.argtypes = [_u64p, C.c_uint64, _u64p]
    L.ora_pcg32_canonical.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, _u32p]
    L.ora_draw_steps_from_state.restype = None
    L.ora_draw_steps_from_state.argtypes = [_u64p, C.POINTER(Jumps), C.c_double, C.c_uint64, _dp, _dp, _dp,
                                            _u32p]
    L.ora_gen_f64_n.argtypes = [C.c_uint64, C.c_uint64, _dp]
    L.ora_gen_range_n.argtypes = [C.c_uint64, C.c_double, C.c_double, C.c_uint64, _dp]
    L.ora_box_muller_n.argtypes = [C.c_uint64, C.c_uint64, _dp]
    L.ora_sample_poisson_n.argtypes = [C.c_uint64, C.c_double, C.c_uint64, _u32p]
    L.ora_qe_variance_step_n.argtypes = [_dp, _dp, C.c_uint64, C.c_double, C.c_double, C.c_double,
                                         C.c_double, _dp]
    L.ora_ncdf_n.argtypes = [_dp, C.c_uint64, _dp]
    L.ora_ncdf.restype = C.c_double
    L.ora_ncdf.argtypes = [C.c_double]
    L.ora_mc_price.restype = C.c_int
    L.ora_mc_price.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.POINTER(Heston),
                               C.POINTER(Jumps), C.POINTER(PayoffS), C.POINTER(Config), C.c_uint64,
                               C.c_int, _dp, C.POINTER(C.c_uint64)]
    L.ora_gen_draws.restype = C.c_int
    L.ora_gen_draws.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_double,
                                C.POINTER(Jumps), _dp, _dp, _dp]
    L.ora_run_paths.restype = C.c_int
    L.ora_run_paths.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.POINTER(Heston),
                                C.POINTER(Jumps), C.POINTER(PayoffS), C.c_uint64, C.c_int, C.c_int,
                                C.c_uint64, _dp, _dp, _dp, _dp, _dp]
    L.ora_heston_price.restype = C.c_double
    L.ora_heston_price.argtypes = [C.c_double] * 5 + [C.POINTER(Heston), C.c_int]
    L.ora_bates_jump_sens.restype = None
    L.ora_bates_jump_sens.argtypes = [C.c_double] * 5 + [C.POINTER(Heston)] + [C.c_double] * 3 + [C.c_int, _dp]
    L.ora_bates_price.restype = C.c_double
    L.ora_bates_price.argtypes = [C.c_double] * 5 + [C.POINTER(Heston)] + [C.c_double] * 3 + [C.c_int]
    L.ora_bsm_greeks.restype = None
    L.ora_bsm_greeks.argtypes = [C.c_double] * 6 + [C.c_int, _dp]
    L.ora_implied_vol.restype = C.c_int
    L.ora_implied_vol.argtypes = [C.c_double] * 5 + [C.c_int, C.c_double, C.POINTER(C.c_double)]
    L.ora_bsm_greeks_many.restype = None
    L.ora_bsm_greeks_many.argtypes = [_dp, C.c_int64, _dp]
    L.ora_implied_vol_many.restype = None
    L.ora_implied_vol_many.argtypes = [_dp, C.c_int64, _dp]
    for nm in ("ora_calib_residuals", "ora_calib_jacobian", "ora_calib_jtj_grad"):
        getattr(L, nm).restype = None
    L.ora_calib_residuals.argtypes = [_dp, C.c_int64, C.c_int, _dp, _dp]
    L.ora_calib_jacobian.argtypes = [_dp, C.c_int64, C.c_int, _dp, _dp]
    L.ora_calib_jtj_grad.argtypes = [_dp, _dp, _dp, C.c_int64, C.c_int, _dp, _dp]
    L.ora_cf_greeks.restype = None
    L.ora_cf_greeks.argtypes = [C.c_double] * 5 + [C.POINTER(Heston), C.c_int] + [C.c_double] * 3 + [C.c_int, _dp]
    L.ora_cf_price_many.restype = None
    L.ora_cf_price_many.argtypes = [_dp, C.c_int64, _dp]
    L.ora_bsm_price.restype = C.c_double
    L.ora_bsm_price.argtypes = [C.c_double] * 6 + [C.c_int]
    L.ora_merton_price.restype = C.c_double
    L.ora_merton_price.argtypes = [C.c_double] * 9 + [C.c_int]
    _lib = L
    return L


# --------------------------------------------------------------- helpers

def heston(v0, kappa, theta, sigma, rho):
    return Heston(v0, kappa, theta, sigma, rho)


def payoff(kind=EUROPEAN, strike=100.0, opt_type=CALL, barrier=0.0, rebate=0.0):
    return PayoffS(kind, opt_type, strike, barrier, rebate)


def jumps(lam=None, mu_j=0.0, sigma_j=0.0):
    if lam is None:
        return Jumps(0, 0.0, 0.0, 0.0)
    return Jumps(1, lam, mu_j, sigma_j)


def mc_price(spot, expiry, rate, div_yield, hp, pay, n_paths, n_steps, seed, antithetic=True,
             scheme=EULER, jump=None, n_threads=1, workers=None):
    """mc_heston / mc_bates (mc.rs:98-113, 148-207) with rayon's thread count
    fixed to ``n_threads``.  Returns (price, std_error, sum, sum_sq, count)."""
    L = lib()
    cfg = Config(n_paths, n_steps, seed, 1 if antithetic else 0, scheme)
    out = np.zeros(4)
    cnt = C.c_uint64(0)
    j = jump if jump is not None else jumps()
    w = n_threads if workers is None else workers
    rc = L.ora_mc_price(spot, expiry, rate, div_yield, C.byref(hp), C.byref(j), C.byref(pay),
                        C.byref(cfg), n_threads, w, out, C.byref(cnt))
    if rc != 0:
        raise RuntimeError("ora_mc_price failed")
    return float(out[0]), float(out[1]), float(out[2]), float(out[3]), int(cnt.value)


def gen_draws(seed, n_pairs, n_steps, expiry, jump=None, n_threads=1):
    """Draws mc_price would consume, as three [n_steps, n_pairs] arrays."""
    L = lib()
    za = np.zeros(n_steps * n_pairs)
    zb = np.zeros(n_steps * n_pairs)
    jp = np.zeros(n_steps * n_pairs)
    j = jump if jump is not None else jumps()
    L.ora_gen_draws(seed, n_pairs, n_steps, n_threads, expiry, C.byref(j), za, zb, jp)
    return (za.reshape(n_steps, n_pairs), zb.reshape(n_steps, n_pairs), jp.reshape(n_steps, n_pairs))


def run_paths(spot, expiry, rate, div_yield, hp, pay, n_steps, scheme, antithetic, za, zb, jump_sum,
              jump=None):
    """run_path (mc.rs:210-278) per pair on given [n_steps, n_pairs] draws.
    Returns payoff [n_pairs, 2] and state [n_pairs, 2, 4] = (ln_s, v, sum, max)."""
    L = lib()
    n_pairs = za.shape[1]
    za = np.ascontiguousarray(za, dtype=np.float64).ravel()
    zb = np.ascontiguousarray(zb, dtype=np.float64).ravel()
    jp = np.ascontiguousarray(jump_sum, dtype=np.float64).ravel()
    pay_out = np.full(n_pairs * 2, np.nan)
    st_out = np.full(n_pairs * 8, np.nan)
    j = jump if jump is not None else jumps()
    L.ora_run_paths(spot, expiry, rate, div_yield, C.byref(hp), C.byref(j), C.byref(pay), n_steps,
                    scheme, 1 if antithetic else 0, n_pairs, za, zb, jp, pay_out, st_out)
    return pay_out.reshape(n_pairs, 2), st_out.reshape(n_pairs, 2, 4)


def heston_price(s, k, t, r, q, hp, opt_type=CALL):
    return lib().ora_heston_price(s, k, t, r, q, C.byref(hp), opt_type)


def bates_price(s, k, t, r, q, hp, lam, mu_j, sigma_j, opt_type=CALL):
    return lib().ora_bates_price(s, k, t, r, q, C.byref(hp), lam, mu_j, sigma_j, opt_type)


GREEK_NAMES = ("price", "delta", "gamma", "vega", "theta", "rho", "vanna", "volga")


def bates_jump_sens(s, k, t, r, q, hp, lam, mu_j, sigma_j, opt_type=CALL):
    """ad.rs:882: (price, d/d lambda, d/d mu_j, d/d sigma_j) of the Bates price."""
    out = np.zeros(4)
    lib().ora_bates_jump_sens(s, k, t, r, q, C.byref(hp), lam, mu_j, sigma_j, opt_type, out)
    return out


def cf_greeks(s, k, t, r, q, hp, jumps=None, opt_type=CALL):
    """greeks.rs bump-and-reprice through heston_price (jumps None) or
    bates_price (jumps = (lambda, mu_j, sigma_j)); a dict of the 8 fields of
    PricingResult."""
    out = np.zeros(8)
    lam, mj, sj = jumps if jumps is not None else (0.0, 0.0, 0.0)
    lib().ora_cf_greeks(s, k, t, r, q, C.byref(hp), int(jumps is not None), lam, mj, sj, opt_type, out)
    return dict(zip(GREEK_NAMES, out))


def bsm_greeks(s, k, t, r, q, vol, opt_type=CALL):
    """bsm.rs:11-45: price and closed-form Greeks, a dict of PricingResult's fields."""
    out = np.zeros(8)
    lib().ora_bsm_greeks(s, k, t, r, q, vol, opt_type, out)
    return dict(zip(GREEK_NAMES, out))


def implied_vol(s, k, t, r, q, opt_type, price):
    """iv.rs:14-117; None where the reference returns None."""
    v = C.c_double()
    return v.value if lib().ora_implied_vol(s, k, t, r, q, opt_type, price, C.byref(v)) else None


def _threaded_rows(fn, args, width_in, width_out, n_threads):
    a = np.ascontiguousarray(args, dtype=np.float64).reshape(-1, width_in)
    n = a.shape[0]
    out = np.zeros((n, width_out)) if width_out > 1 else np.zeros(n)
    bounds = np.linspace(0, n, max(1, n_threads) + 1).astype(int)

    def run(k):
        lo, hi = int(bounds[k]), int(bounds[k + 1])
        if hi > lo:
            fn(a[lo:hi], hi - lo, out[lo:hi])

    import concurrent.futures
    with concurrent.futures.ThreadPoolExecutor(max(1, n_threads)) as ex:
        list(ex.map(run, range(len(bounds) - 1)))
    return out


def bsm_greeks_many(args, n_threads=1):
    """Rows (spot, strike, expiry, rate, div_yield, vol, opt_type) -> n x 8 price and Greeks."""
    return _threaded_rows(lib().ora_bsm_greeks_many, args, 7, 8, n_threads)


def implied_vol_many(args, n_threads=1):
    """Rows (spot, strike, expiry, rate, div_yield, opt_type, price) -> vols (NaN: none)."""
    return _threaded_rows(lib().ora_implied_vol_many, args, 7, 1, n_threads)


def calib_eval(quotes, model, params):
    """calibration.rs residuals (573-587), FD Jacobian (613-652) and J'J / J'r
    (655-667).  quotes: n x 8 rows (spot, strike, expiry, rate, div_yield,
    opt_type, iv_market, weight); model 0 Heston, 1 Bates (every column
    differenced).  Returns (residuals, jacobian n x dim, jtj, grad)."""
    q = np.ascontiguousarray(quotes, dtype=np.float64).reshape(-1, 8)
    n, dim = q.shape[0], 8 if model else 5
    p = np.ascontiguousarray(params, dtype=np.float64)
    res, jac = np.zeros(n), np.zeros((n, dim))
    jtj, grad = np.zeros((dim, dim)), np.zeros(dim)
    L = lib()
    L.ora_calib_residuals(q, n, model, p, res)
    L.ora_calib_jacobian(q, n, model, p, jac)
    L.ora_calib_jtj_grad(jac, res, np.ascontiguousarray(q[:, 7]), n, dim, jtj, grad)
    return res, jac, jtj, grad


def cf_price_many(args, n_threads=1):
    """Analytic prices of the rows of args (n x 15: spot, strike, expiry,
    rate, div_yield, v0, kappa, theta, sigma, rho, lambda, mu_j, sigma_j,
    model 0 Heston / 1 Bates, opt_type), split over n_threads host threads
    (the C loop runs without the GIL)."""
    import concurrent.futures

    a = np.ascontiguousarray(args, dtype=np.float64).reshape(-1, 15)
    n = a.shape[0]
    out = np.zeros(n)
    bounds = np.linspace(0, n, max(1, n_threads) + 1).astype(int)

    def run(k):
        lo, hi = int(bounds[k]), int(bounds[k + 1])
        if hi > lo:
            lib().ora_cf_price_many(a[lo:hi], hi - lo, out[lo:hi])

    with concurrent.futures.ThreadPoolExecutor(max(1, n_threads)) as ex:
        list(ex.map(run, range(len(bounds) - 1)))
    return out


def bsm_price(s, k, t, r, q, vol, opt_type=CALL):
    return lib().ora_bsm_price(s, k, t, r, q, vol, opt_type)


def merton_price(s, k, t, r, q, var, lam, mu_j, sigma_j, opt_type=CALL):
    return lib().ora_merton_price(s, k, t, r, q, var, lam, mu_j, sigma_j, opt_type)


def qe_variance_step(v_t, z, kappa, theta, sigma, dt):
    v_t = np.ascontiguousarray(v_t, dtype=np.float64)
    z = np.ascontiguousarray(z, dtype=np.float64)
    v_t, z = np.broadcast_arrays(v_t, z)
    v_t = np.ascontiguousarray(v_t)
    z = np.ascontiguousarray(z)
    out = np.zeros(v_t.size)
    lib().ora_qe_variance_step_n(v_t.ravel(), z.ravel(), v_t.size, kappa, theta, sigma, dt, out)
    return out.reshape(v_t.shape)


def ncdf(x):
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    out = np.zeros(x.size)
    lib().ora_ncdf_n(x, x.size, out)
    return out


# ------------------------------------------- Philox4x32-10 stream oracle

M32 = 0xFFFFFFFF
PHILOX_M0, PHILOX_M1 = 0xD2511F53, 0xCD9E8D57
PHILOX_W0, PHILOX_W1 = 0x9E3779B9, 0xBB67AE85
TAG_NORMAL, TAG_JUMP = 0, 2
GATE_BITS = 9


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Philox4x32 with 10 rounds (Salmon et al. 2011).  Counters are uint64
    arrays holding 32-bit values; the key is two Python ints."""
    c0, c1, c2, c3 = (np.asarray(c, dtype=np.uint64) & np.uint64(M32) for c in (c0, c1, c2, c3))
    k0 = int(k0) & M32
    k1 = int(k1) & M32
    m0, m1, mask = np.uint64(PHILOX_M0), np.uint64(PHILOX_M1), np.uint64(M32)
    for _ in range(10):
        p0 = m0 * c0
        p1 = m1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & mask
        hi1, lo1 = p1 >> np.uint64(32), p1 & mask
        c0, c1, c2, c3 = (hi1 ^ c1 ^ np.uint64(k0), lo1, hi0 ^ c3 ^ np.uint64(k1), lo0)
        k0 = (k0 + PHILOX_W0) & M32
        k1 = (k1 + PHILOX_W1) & M32
    return c0, c1, c2, c3


def poisson_threshold0(mu):
    """Integer threshold T with (w < T) <=> ((w + 0.5) 2^-32 < e^-mu)."""
    f0 = math.exp(-mu)
    return int(math.ceil(f0 * 4294967296.0 - 0.5))


def poisson_from_word(w, mu, kmax=1024):
    """Inverse-CDF Poisson count of the uniform (w + 0.5) 2^-32 (FP64 CDF)."""
    out = np.zeros(np.shape(w), dtype=np.int64)
    if mu <= 0.0:
        return out
    thr = poisson_threshold0(mu)
    e0 = math.exp(-mu)
    flat_w = np.asarray(w, dtype=np.uint64).ravel()
    flat = out.ravel()
    for idx in np.nonzero(flat_w >= np.uint64(thr))[0]:
        u = (float(flat_w[idx]) + 0.5) * 2.0 ** -32
        k, p, F = 0, e0, e0
        while u >= F and k < kmax:
            k += 1
            p = p * mu / k
            F = F + p
        flat[idx] = k
    return flat.reshape(np.shape(w))


def _box_muller(wa, wb):
    u1 = (wa.astype(np.float64) + 0.5) * 2.0 ** -32
    u2 = (wb >> np.uint64(GATE_BITS)).astype(np.float64) * 2.0 ** -23
    r = np.sqrt(-2.0 * np.log(u1))
    return r * np.cos(2.0 * np.pi * u2), r * np.sin(2.0 * np.pi * u2)


def philox_draws(seed, pair_ids, n_steps, dt=None, lam=0.0, mu_j=0.0, sigma_j=0.0):
    """The GPU engine's Philox-mode draws, FP64, for the given pair indices.

    Counter = (call, pair lo, pair hi, tag), key = (seed lo, seed hi).
    Normals of step s: call s >> 1 with tag 0, words (wa, wb) = (0,1) for
    even s and (2,3) for odd s, mapped by Box-Muller with u1 = (wa + 1/2)
    2^-32 and the 23-bit angle u2 = (wb >> 9) 2^-23.  Bates (lam > 0) adds
    the Poisson count of step s, by inverse CDF of mean lam*dt of the 32-bit
    word W = (h << 23) | l: h is the 9 low bits of wb; when h >= T >> 23 (T
    the count-0 threshold) l is the 23 high bits of word 2 of call s with
    tag 2, otherwise the count is 0 whatever l is.  A non-zero count takes
    its jump normal from the Box-Muller cosine output of words (0,1) of that
    tag-2 call.  Returns za, zb, jump_sum, n as [n_steps, n_pairs].
    """
    pair_ids = np.asarray(pair_ids, dtype=np.uint64)
    npairs = pair_ids.size
    k0, k1 = int(seed) & M32, (int(seed) >> 32) & M32
    plo = pair_ids & np.uint64(M32)
    phi = pair_ids >> np.uint64(32)
    za = np.zeros((n_steps, npairs))
    zb = np.zeros((n_steps, npairs))
    js = np.zeros((n_steps, npairs))
    nn = np.zeros((n_steps, npairs), dtype=np.int64)
    bates = lam > 0.0

    def call(c, tag):
        return philox4x32_10(np.full(npairs, c, dtype=np.uint64), plo, phi, np.full(npairs, tag, dtype=np.uint64),
                             k0, k1)

    for s in range(n_steps):
        w = call(s >> 1, TAG_NORMAL)
        wa, wb = (w[0], w[1]) if (s & 1) == 0 else (w[2], w[3])
        za[s], zb[s] = _box_muller(wa, wb)
        if bates:
            mu = lam * dt
            h = wb & np.uint64((1 << GATE_BITS) - 1)
            gate = min(poisson_threshold0(mu), M32) >> (32 - GATE_BITS)
            maybe = h >= np.uint64(gate)
            if np.any(maybe):
                wj = call(s, TAG_JUMP)
                full = (h << np.uint64(32 - GATE_BITS)) | (wj[2] >> np.uint64(GATE_BITS))
                n = np.where(maybe, poisson_from_word(full, mu), 0)
                zj, _ = _box_muller(wj[0], wj[1])
                js[s] = np.where(n > 0, mu_j * n + sigma_j * np.sqrt(n) * zj, 0.0)
                nn[s] = n
    return za, zb, js, nn
