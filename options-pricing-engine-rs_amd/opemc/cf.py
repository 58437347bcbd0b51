"""Python mirror of the reference's analytic pricers over the GPU kernel.

``heston_price`` (heston.rs:61-72), ``bates_price`` (bates.rs:29-39),
``batch_heston`` / ``batch_bates`` (batch.rs:23-27, 38-42), ``BatesParams``
(types.rs:67-72) and ``OptionContract`` (types.rs:17-25), through the C ABI of
``include/ope_cf.h``.  ``cf_price_options`` prices options with a model each
(one launch) and can report the quadrature panels and the GPU time.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from typing import Optional, Sequence, Tuple

import numpy as np

from . import HestonParams, OptionType, _Heston, _check, lib

CF_HESTON = 0
CF_BATES = 1


class _Bates(C.Structure):
    _fields_ = [("heston", _Heston), ("lambda_", C.c_double), ("mu_j", C.c_double), ("sigma_j", C.c_double)]


class _Contract(C.Structure):
    _fields_ = [("spot", C.c_double), ("strike", C.c_double), ("expiry", C.c_double), ("rate", C.c_double),
                ("div_yield", C.c_double), ("vol", C.c_double), ("opt_type", C.c_int32),
                ("reserved_", C.c_int32)]


class _CfOption(C.Structure):
    _fields_ = [("contract", _Contract), ("params", _Bates), ("model", C.c_int32), ("reserved_", C.c_int32)]


class _PricingResult(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("price", "delta", "gamma", "vega", "theta", "rho", "vanna", "volga")]


@dataclasses.dataclass(frozen=True)
class PricingResult:
    """types.rs:29-39"""
    price: float
    delta: float
    gamma: float
    vega: float
    theta: float
    rho: float
    vanna: float
    volga: float

    @classmethod
    def _from(cls, r):
        return cls(r.price, r.delta, r.gamma, r.vega, r.theta, r.rho, r.vanna, r.volga)


def _cf_lib():
    L = lib()
    if not getattr(L, "_cf_ready", False):
        d = C.c_double
        L.ope_heston_price.argtypes = [d, d, d, d, d, C.POINTER(_Heston), C.c_int32, C.POINTER(d)]
        L.ope_bates_price.argtypes = [d, d, d, d, d, C.POINTER(_Bates), C.c_int32, C.POINTER(d)]
        L.ope_batch_heston.argtypes = [C.POINTER(_Contract), C.c_size_t, C.POINTER(_Heston), C.POINTER(d)]
        L.ope_batch_bates.argtypes = [C.POINTER(_Contract), C.c_size_t, C.POINTER(_Bates), C.POINTER(d)]
        L.ope_cf_price_options.argtypes = [C.POINTER(_CfOption), C.c_size_t, C.POINTER(d),
                                           C.POINTER(C.c_uint32), C.POINTER(d)]
        R = C.POINTER(_PricingResult)
        L.ope_heston_price_and_greeks.argtypes = [d, d, d, d, d, C.POINTER(_Heston), C.c_int32, R]
        L.ope_bates_price_and_greeks.argtypes = [d, d, d, d, d, C.POINTER(_Bates), C.c_int32, R]
        L.ope_batch_heston_greeks.argtypes = [C.POINTER(_Contract), C.c_size_t, C.POINTER(_Heston), R]
        L.ope_batch_bates_greeks.argtypes = [C.POINTER(_Contract), C.c_size_t, C.POINTER(_Bates), R]
        K = C.POINTER(_Contract)
        L.ope_bsm_price_and_greeks.argtypes = [K, R]
        L.ope_bsm_price.argtypes = [K, C.POINTER(d)]
        L.ope_implied_vol.argtypes = [K, d, C.POINTER(d), C.POINTER(C.c_uint8)]
        L.ope_batch_bsm.argtypes = [K, C.c_size_t, R]
        L.ope_batch_bsm_price.argtypes = [K, C.c_size_t, C.POINTER(d)]
        L.ope_batch_implied_vol.argtypes = [K, C.POINTER(d), C.c_size_t, C.POINTER(d), C.POINTER(C.c_uint8)]
        L.ope_batch_bsm_ex.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.POINTER(d)]
        L.ope_batch_implied_vol_ex.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p,
                                               C.POINTER(d)]
        for n in ("ope_heston_price", "ope_bates_price", "ope_batch_heston", "ope_batch_bates",
                  "ope_cf_price_options", "ope_heston_price_and_greeks", "ope_bates_price_and_greeks",
                  "ope_batch_heston_greeks", "ope_batch_bates_greeks", "ope_bsm_price_and_greeks",
                  "ope_bsm_price", "ope_implied_vol", "ope_batch_bsm", "ope_batch_bsm_price",
                  "ope_batch_implied_vol", "ope_batch_bsm_ex", "ope_batch_implied_vol_ex"):
            getattr(L, n).restype = C.c_int
        L._cf_ready = True
    return L


@dataclasses.dataclass(frozen=True)
class BatesParams:
    """types.rs:67-72"""
    heston: HestonParams
    lambda_: float
    mu_j: float
    sigma_j: float

    def _c(self):
        return _Bates(self.heston._c(), self.lambda_, self.mu_j, self.sigma_j)


@dataclasses.dataclass(frozen=True)
class OptionContract:
    """types.rs:17-25 (vol is the BSM volatility; the CF pricers ignore it)."""
    spot: float
    strike: float
    expiry: float
    rate: float
    div_yield: float
    vol: float = 0.0
    opt_type: OptionType = OptionType.Call

    def _c(self):
        return _Contract(self.spot, self.strike, self.expiry, self.rate, self.div_yield, self.vol,
                         int(self.opt_type), 0)


def heston_price(spot, strike, expiry, rate, div_yield, params: HestonParams,
                 opt_type: OptionType = OptionType.Call) -> float:
    out = C.c_double()
    hp = params._c()
    _check(_cf_lib().ope_heston_price(spot, strike, expiry, rate, div_yield, C.byref(hp), int(opt_type),
                                      C.byref(out)))
    return out.value


def bates_price(spot, strike, expiry, rate, div_yield, params: BatesParams,
                opt_type: OptionType = OptionType.Call) -> float:
    out = C.c_double()
    bp = params._c()
    _check(_cf_lib().ope_bates_price(spot, strike, expiry, rate, div_yield, C.byref(bp), int(opt_type),
                                     C.byref(out)))
    return out.value


def _contracts(contracts: Sequence[OptionContract]):
    arr = (_Contract * max(1, len(contracts)))()
    for i, c in enumerate(contracts):
        arr[i] = c._c()
    return arr


def batch_heston(contracts: Sequence[OptionContract], params: HestonParams) -> np.ndarray:
    out = np.zeros(len(contracts))
    hp = params._c()
    _check(_cf_lib().ope_batch_heston(_contracts(contracts), len(contracts), C.byref(hp),
                                      out.ctypes.data_as(C.POINTER(C.c_double))))
    return out


def batch_bates(contracts: Sequence[OptionContract], params: BatesParams) -> np.ndarray:
    out = np.zeros(len(contracts))
    bp = params._c()
    _check(_cf_lib().ope_batch_bates(_contracts(contracts), len(contracts), C.byref(bp),
                                     out.ctypes.data_as(C.POINTER(C.c_double))))
    return out


def heston_price_and_greeks(spot, strike, expiry, rate, div_yield, params: HestonParams,
                            opt_type: OptionType = OptionType.Call) -> PricingResult:
    """heston.rs:74-80, bump-and-reprice (greeks.rs:19-82) on the GPU."""
    out = _PricingResult()
    hp = params._c()
    _check(_cf_lib().ope_heston_price_and_greeks(spot, strike, expiry, rate, div_yield, C.byref(hp), int(opt_type),
                                                 C.byref(out)))
    return PricingResult._from(out)


def bates_price_and_greeks(spot, strike, expiry, rate, div_yield, params: BatesParams,
                           opt_type: OptionType = OptionType.Call) -> PricingResult:
    """bates.rs:41-47"""
    out = _PricingResult()
    bp = params._c()
    _check(_cf_lib().ope_bates_price_and_greeks(spot, strike, expiry, rate, div_yield, C.byref(bp), int(opt_type),
                                                C.byref(out)))
    return PricingResult._from(out)


def batch_heston_greeks(contracts: Sequence[OptionContract], params: HestonParams):
    """batch.rs:32-36: one launch for every bumped price of the chain."""
    out = (_PricingResult * max(1, len(contracts)))()
    hp = params._c()
    _check(_cf_lib().ope_batch_heston_greeks(_contracts(contracts), len(contracts), C.byref(hp), out))
    return [PricingResult._from(out[i]) for i in range(len(contracts))]


def batch_bates_greeks(contracts: Sequence[OptionContract], params: BatesParams):
    """batch.rs:44-48"""
    out = (_PricingResult * max(1, len(contracts)))()
    bp = params._c()
    _check(_cf_lib().ope_batch_bates_greeks(_contracts(contracts), len(contracts), C.byref(bp), out))
    return [PricingResult._from(out[i]) for i in range(len(contracts))]


@dataclasses.dataclass(frozen=True)
class CfOption:
    """One option with its own model: params is HestonParams or BatesParams."""
    contract: OptionContract
    params: object

    def _c(self):
        if isinstance(self.params, BatesParams):
            return _CfOption(self.contract._c(), self.params._c(), CF_BATES, 0)
        return _CfOption(self.contract._c(), _Bates(self.params._c(), 0.0, 0.0, 0.0), CF_HESTON, 0)


# numpy layout of ope_cf_option, for building large batches without a
# Python object per option
CF_DTYPE = np.dtype([("spot", "f8"), ("strike", "f8"), ("expiry", "f8"), ("rate", "f8"), ("div_yield", "f8"),
                     ("vol", "f8"), ("opt_type", "i4"), ("reserved0", "i4"), ("v0", "f8"), ("kappa", "f8"),
                     ("theta", "f8"), ("sigma", "f8"), ("rho", "f8"), ("lambda_", "f8"), ("mu_j", "f8"),
                     ("sigma_j", "f8"), ("model", "i4"), ("reserved1", "i4")])
assert CF_DTYPE.itemsize == C.sizeof(_CfOption)


def cf_price_options(options, with_panels=False, with_kernel_ms=False):
    """Prices of n options in one launch.  options: a sequence of CfOption,
    or a numpy array of CF_DTYPE.  Returns the prices, plus the (n, 2) panel
    counts of the P1 / P2 integrals and the GPU time in ms when asked."""
    if isinstance(options, np.ndarray):
        if options.dtype != CF_DTYPE:
            raise TypeError("options array must have dtype CF_DTYPE")
        n = len(options)
        packed = np.ascontiguousarray(options)
        arr = packed.ctypes.data_as(C.POINTER(_CfOption))
    else:
        n = len(options)
        arr = (_CfOption * max(1, n))()
        for i, o in enumerate(options):
            arr[i] = o._c()
    out = np.zeros(n)
    panels = np.zeros((n, 2), dtype=np.uint32) if with_panels else None
    kms = C.c_double()
    _check(_cf_lib().ope_cf_price_options(arr, n, out.ctypes.data_as(C.POINTER(C.c_double)),
                                          panels.ctypes.data_as(C.POINTER(C.c_uint32)) if with_panels else None,
                                          C.byref(kms)))
    res: Tuple = (out,)
    if with_panels:
        res += (panels,)
    if with_kernel_ms:
        res += (kms.value,)
    return res[0] if len(res) == 1 else res


# ------------------------------------------------ Black-Scholes-Merton, IV

# numpy layout of ope_option_contract (types.rs:17-25), for large chains
CONTRACT_DTYPE = np.dtype([("spot", "f8"), ("strike", "f8"), ("expiry", "f8"), ("rate", "f8"),
                           ("div_yield", "f8"), ("vol", "f8"), ("opt_type", "i4"), ("reserved", "i4")])
GREEKS_DTYPE = np.dtype([(n, "f8") for n in ("price", "delta", "gamma", "vega", "theta", "rho", "vanna", "volga")])
assert CONTRACT_DTYPE.itemsize == C.sizeof(_Contract) and GREEKS_DTYPE.itemsize == C.sizeof(_PricingResult)


def bsm_price_and_greeks(contract: OptionContract) -> PricingResult:
    """bsm.rs:11-35"""
    out = _PricingResult()
    k = contract._c()
    _check(_cf_lib().ope_bsm_price_and_greeks(C.byref(k), C.byref(out)))
    return PricingResult._from(out)


def bsm_price(contract: OptionContract) -> float:
    """bsm.rs:49-54"""
    out = C.c_double()
    k = contract._c()
    _check(_cf_lib().ope_bsm_price(C.byref(k), C.byref(out)))
    return out.value


def implied_vol(contract: OptionContract, market_price: float) -> Optional[float]:
    """iv.rs:14-20: None where no volatility is found."""
    v, f = C.c_double(), C.c_uint8()
    k = contract._c()
    _check(_cf_lib().ope_implied_vol(C.byref(k), market_price, C.byref(v), C.byref(f)))
    return v.value if f.value else None


def batch_bsm(contracts: Sequence[OptionContract]):
    """batch.rs:13-15"""
    out = (_PricingResult * max(1, len(contracts)))()
    _check(_cf_lib().ope_batch_bsm(_contracts(contracts), len(contracts), out))
    return [PricingResult._from(out[i]) for i in range(len(contracts))]


def batch_bsm_price(contracts: Sequence[OptionContract]) -> np.ndarray:
    """batch.rs:18-20"""
    out = np.zeros(len(contracts))
    _check(_cf_lib().ope_batch_bsm_price(_contracts(contracts), len(contracts),
                                         out.ctypes.data_as(C.POINTER(C.c_double))))
    return out


def batch_implied_vol(contracts: Sequence[OptionContract], market_prices: Sequence[float]):
    """batch.rs:51-56: a list with None where the solver found no volatility."""
    n = len(contracts)
    if len(market_prices) != n:
        raise ValueError("one market price per contract")
    mkt = np.ascontiguousarray(market_prices, dtype=np.float64)
    vol = np.zeros(n)
    found = np.zeros(n, dtype=np.uint8)
    _check(_cf_lib().ope_batch_implied_vol(_contracts(contracts), mkt.ctypes.data_as(C.POINTER(C.c_double)), n,
                                           vol.ctypes.data_as(C.POINTER(C.c_double)),
                                           found.ctypes.data_as(C.POINTER(C.c_uint8))))
    return [float(v) if f else None for v, f in zip(vol, found)]


def bsm_chain(contracts: np.ndarray, greeks=True):
    """A CONTRACT_DTYPE chain in one launch: (GREEKS_DTYPE records or prices, kernel ms)."""
    c = np.ascontiguousarray(contracts)
    if c.dtype != CONTRACT_DTYPE:
        raise TypeError("contracts must have dtype CONTRACT_DTYPE")
    out = np.zeros(len(c), dtype=GREEKS_DTYPE) if greeks else np.zeros(len(c))
    kms = C.c_double()
    _check(_cf_lib().ope_batch_bsm_ex(c.ctypes.data, len(c), None if greeks else out.ctypes.data,
                                      out.ctypes.data if greeks else None, C.byref(kms)))
    return out, kms.value


def implied_vol_chain(contracts: np.ndarray, market_prices: np.ndarray):
    """A CONTRACT_DTYPE chain in one launch: (vols with NaN where none, found flags, kernel ms)."""
    c = np.ascontiguousarray(contracts)
    if c.dtype != CONTRACT_DTYPE:
        raise TypeError("contracts must have dtype CONTRACT_DTYPE")
    mkt = np.ascontiguousarray(market_prices, dtype=np.float64)
    vol = np.zeros(len(c))
    found = np.zeros(len(c), dtype=np.uint8)
    kms = C.c_double()
    _check(_cf_lib().ope_batch_implied_vol_ex(c.ctypes.data, mkt.ctypes.data, len(c), vol.ctypes.data,
                                              found.ctypes.data, C.byref(kms)))
    return vol, found, kms.value


# ------------------------------------------------------------ calibration

CALIB_DTYPE = np.dtype([("contract", CONTRACT_DTYPE), ("iv_market", "f8"), ("weight", "f8")])


def _calib_lib():
    L = _cf_lib()
    if not getattr(L, "_calib_ready", False):
        P = C.c_void_p
        L.ope_calib_residuals.argtypes = [P, C.c_size_t, C.c_int32, P, P]
        L.ope_calib_jacobian.argtypes = [P, C.c_size_t, C.c_int32, P, P, P, P, P]
        L.ope_calib_residuals.restype = C.c_int
        L.ope_calib_jacobian.restype = C.c_int
        L._calib_ready = True
    return L


def _calib_params(params):
    if isinstance(params, BatesParams):
        h = params.heston
        return CF_BATES, np.array([h.v0, h.kappa, h.theta, h.sigma, h.rho, params.lambda_, params.mu_j,
                                   params.sigma_j])
    if isinstance(params, HestonParams):
        return CF_HESTON, np.array([params.v0, params.kappa, params.theta, params.sigma, params.rho])
    raise TypeError("params must be HestonParams or BatesParams")


def calib_residuals(quotes: np.ndarray, params) -> np.ndarray:
    """calibration.rs:573-587 on the GPU: iv(model price) - iv_market per quote
    (0 where no volatility is found).  quotes: CALIB_DTYPE records."""
    q = np.ascontiguousarray(quotes)
    if q.dtype != CALIB_DTYPE:
        raise TypeError("quotes must have dtype CALIB_DTYPE")
    model, p = _calib_params(params)
    res = np.zeros(len(q))
    _check(_calib_lib().ope_calib_residuals(q.ctypes.data, len(q), model, p.ctypes.data, res.ctypes.data))
    return res


def calib_jacobian(quotes: np.ndarray, params):
    """calibration.rs:613-667 on the GPU: (residuals, Jacobian n x dim, J'J, J'r)."""
    q = np.ascontiguousarray(quotes)
    if q.dtype != CALIB_DTYPE:
        raise TypeError("quotes must have dtype CALIB_DTYPE")
    model, p = _calib_params(params)
    n, dim = len(q), len(p)
    res, jac = np.zeros(n), np.zeros((n, dim))
    jtj, grad = np.zeros((dim, dim)), np.zeros(dim)
    _check(_calib_lib().ope_calib_jacobian(q.ctypes.data, n, model, p.ctypes.data, res.ctypes.data,
                                           jac.ctypes.data, jtj.ctypes.data, grad.ctypes.data))
    return res, jac, jtj, grad
