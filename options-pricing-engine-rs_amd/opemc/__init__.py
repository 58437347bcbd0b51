"""Python mirror of the reference's Monte Carlo API over the HIP engine.

The names, argument order and meaning follow ``src/mc.rs`` of the reference
(``mc_heston`` mc.rs:98-103, ``mc_bates`` mc.rs:107-113, ``McConfig`` with its
``Default`` mc.rs:48-64, ``Payoff`` mc.rs:66-72, ``VarianceScheme`` mc.rs:34-46,
``McResult`` mc.rs:28-32) and ``src/types.rs`` (``HestonParams`` 47-55,
``OptionType`` 5-13).  Every call goes through the C ABI of
``lib/libopemc.so`` (``include/ope_mc.h``) to the GPU kernels.  There is no
CPU fallback: a missing library raises at import of the first call, and a
missing GPU raises ``OpeError`` with status ``OPE_ERR_NO_DEVICE``.

The reference's functions are infallible (NaN on degenerate input); here a
non-zero status raises ``OpeError``, the Python counterpart of the Rust
binding's panic.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import enum
import os
from typing import NamedTuple, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libopemc.so")

OPE_OK = 0
OPE_ERR_INVALID_ARGUMENT = -1
OPE_ERR_NO_DEVICE = -2
OPE_ERR_HIP = -3
OPE_ERR_OUT_OF_MEMORY = -4
OPE_ERR_OVER_BUDGET = -5
OPE_ERR_NOT_FOUND = -6
OPE_ERR_SHUTDOWN = -7
OPE_ERR_CANCELLED = -8
OPE_ERR_DEADLINE = -9
OPE_ERR_TIMEOUT = -10

FLAG_BARRIER_FP64 = 1

SERVER_STOP_ALL_PATHS = 0
SERVER_STOP_TARGET = 1
SERVER_STOP_DEADLINE = 2

PRECISION_FP64 = 0
PRECISION_FP32_PATH = 1
RNG_PHILOX = 0
RNG_REPLAY = 1


class OpeError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"[{status}] {message}")
        self.status = status


# --------------------------------------------------------------- C types

class _Heston(C.Structure):
    _fields_ = [("v0", C.c_double), ("kappa", C.c_double), ("theta", C.c_double),
                ("sigma", C.c_double), ("rho", C.c_double)]


class _Payoff(C.Structure):
    _fields_ = [("kind", C.c_int32), ("opt_type", C.c_int32), ("strike", C.c_double),
                ("barrier", C.c_double), ("rebate", C.c_double)]


class _Config(C.Structure):
    _fields_ = [("n_paths", C.c_uint64), ("n_steps", C.c_uint64), ("seed", C.c_uint64),
                ("antithetic", C.c_uint8), ("reserved_", C.c_uint8 * 3), ("scheme", C.c_int32)]


class _Result(C.Structure):
    _fields_ = [("price", C.c_double), ("std_error", C.c_double)]


class _Stats(C.Structure):
    _fields_ = [("sum", C.c_double), ("sum_sq", C.c_double), ("count", C.c_uint64),
                ("path_steps", C.c_uint64)]


class _Replay(C.Structure):
    _fields_ = [("za", C.POINTER(C.c_double)), ("zb", C.POINTER(C.c_double)),
                ("jump_sum", C.POINTER(C.c_double)), ("n_pairs", C.c_uint64), ("n_steps", C.c_uint64)]


class _Options(C.Structure):
    _fields_ = [("precision", C.c_int32), ("rng", C.c_int32), ("replay", C.POINTER(_Replay)),
                ("path_payoffs", C.POINTER(C.c_double)), ("path_states", C.POINTER(C.c_double)),
                ("stats", C.POINTER(_Stats)), ("device_ms", C.POINTER(C.c_double)),
                ("kernel_ms", C.POINTER(C.c_double)), ("flags", C.c_uint32), ("reserved_", C.c_uint32)]


class _Request(C.Structure):
    _fields_ = [("spot", C.c_double), ("expiry", C.c_double), ("rate", C.c_double),
                ("div_yield", C.c_double), ("heston", _Heston), ("lambda_", C.c_double),
                ("mu_j", C.c_double), ("sigma_j", C.c_double), ("payoff", _Payoff),
                ("cfg", _Config)]


def _flat_dtype(struct) -> np.dtype:
    """A numpy dtype with the byte layout of a ctypes Structure, nested
    structures flattened to dotted names and byte arrays (reserved fields)
    left as zero padding, so a batch of structures fills from Python tuples
    in one assignment."""
    names, formats, offsets = [], [], []

    def walk(st, base, prefix):
        for name, ct, *_ in st._fields_:
            off = base + getattr(st, name).offset
            if issubclass(ct, C.Structure):
                walk(ct, off, prefix + name + ".")
            elif not issubclass(ct, C.Array):
                names.append(prefix + name)
                formats.append(np.dtype(ct))
                offsets.append(off)

    walk(struct, 0, "")
    return np.dtype({"names": names, "formats": formats, "offsets": offsets, "itemsize": C.sizeof(struct)})


_REQUEST_NP = _flat_dtype(_Request)
_RESULT_NP = _flat_dtype(_Result)
_STATS_NP = _flat_dtype(_Stats)


class _DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("arch", C.c_char * 32), ("compute_units", C.c_int32),
                ("clock_mhz", C.c_int32), ("total_mem_bytes", C.c_uint64)]


_lib = None


def lib():
    """Load the engine library.  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"HIP engine library not built: {LIB_PATH} "
                          "(run `python options-pricing-engine-rs_amd/build.py`)")
    L = C.CDLL(LIB_PATH)
    d = C.c_double
    L.ope_mc_config_default.argtypes = [C.POINTER(_Config)]
    L.ope_mc_config_default.restype = None
    L.ope_mc_heston.argtypes = [d, d, d, d, C.POINTER(_Heston), C.POINTER(_Payoff), C.POINTER(_Config),
                                C.POINTER(_Result)]
    L.ope_mc_bates.argtypes = [d, d, d, d, C.POINTER(_Heston), d, d, d, C.POINTER(_Payoff),
                               C.POINTER(_Config), C.POINTER(_Result)]
    L.ope_mc_heston_ex.argtypes = [d, d, d, d, C.POINTER(_Heston), C.POINTER(_Payoff), C.POINTER(_Config),
                                   C.POINTER(_Options), C.POINTER(_Result)]
    L.ope_mc_bates_ex.argtypes = [d, d, d, d, C.POINTER(_Heston), d, d, d, C.POINTER(_Payoff),
                                  C.POINTER(_Config), C.POINTER(_Options), C.POINTER(_Result)]
    L.ope_mc_surface.argtypes = [d, d, d, C.POINTER(_Heston), d, d, d, C.c_int32, C.POINTER(C.c_double),
                                 C.c_uint32, C.POINTER(C.c_double), C.c_uint32, C.POINTER(_Config),
                                 C.POINTER(_Options), C.POINTER(_Result)]
    L.ope_mc_batch.argtypes = [C.POINTER(_Request), C.c_uint32, C.POINTER(_Options), C.POINTER(_Result)]
    L.ope_get_device_info.argtypes = [C.POINTER(_DeviceInfo)]
    L.ope_status_string.argtypes = [C.c_int]
    L.ope_status_string.restype = C.c_char_p
    L.ope_last_error_message.argtypes = []
    L.ope_last_error_message.restype = C.c_char_p
    L.ope_debug_philox.argtypes = [C.c_uint64, C.POINTER(C.c_uint32), C.c_uint32, C.POINTER(C.c_uint32)]
    L.ope_debug_draws.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, d, d, d, d, C.c_int32,
                                  C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                  C.POINTER(C.c_uint32)]
    L.ope_debug_fp64_math.argtypes = [C.c_int32, C.POINTER(C.c_double), C.c_uint32, C.POINTER(C.c_double)]
    for name in ("ope_mc_heston", "ope_mc_bates", "ope_mc_heston_ex", "ope_mc_bates_ex", "ope_mc_surface",
                 "ope_mc_batch", "ope_get_device_info", "ope_debug_philox", "ope_debug_draws",
                 "ope_debug_fp64_math"):
        getattr(L, name).restype = C.c_int
    _lib = L
    return L


def _check(rc: int):
    if rc != OPE_OK:
        L = lib()
        msg = L.ope_last_error_message().decode(errors="replace") or L.ope_status_string(rc).decode()
        raise OpeError(rc, msg)


# ------------------------------------------------- reference-shaped API

class OptionType(enum.IntEnum):
    """types.rs:5-13"""
    Call = 0
    Put = 1

    def sign(self) -> float:
        return 1.0 if self == OptionType.Call else -1.0


class VarianceScheme(enum.IntEnum):
    """mc.rs:34-46"""
    FullTruncationEuler = 0
    QuadraticExponential = 1


@dataclasses.dataclass(frozen=True)
class HestonParams:
    """types.rs:47-55 (v0 is a variance)."""
    v0: float
    kappa: float
    theta: float
    sigma: float
    rho: float

    def feller_ok(self) -> bool:
        """types.rs:57-63"""
        return 2.0 * self.kappa * self.theta > self.sigma * self.sigma

    def _c(self):
        return _Heston(self.v0, self.kappa, self.theta, self.sigma, self.rho)


class Payoff:
    """mc.rs:66-72: Payoff.European / Payoff.AsianArithmetic / Payoff.UpAndOut."""
    EUROPEAN, ASIAN_ARITHMETIC, UP_AND_OUT = 0, 1, 2

    def __init__(self, kind, strike, opt_type=OptionType.Call, barrier=0.0, rebate=0.0):
        self.kind = int(kind)
        self.strike = float(strike)
        self.opt_type = OptionType(opt_type)
        self.barrier = float(barrier)
        self.rebate = float(rebate)

    @classmethod
    def European(cls, strike, opt_type=OptionType.Call):
        return cls(cls.EUROPEAN, strike, opt_type)

    @classmethod
    def AsianArithmetic(cls, strike, opt_type=OptionType.Call):
        return cls(cls.ASIAN_ARITHMETIC, strike, opt_type)

    @classmethod
    def UpAndOut(cls, strike, barrier, rebate=0.0, opt_type=OptionType.Call):
        return cls(cls.UP_AND_OUT, strike, opt_type, barrier, rebate)

    def _c(self):
        return _Payoff(self.kind, int(self.opt_type), self.strike, self.barrier, self.rebate)

    def __repr__(self):
        names = {0: "European", 1: "AsianArithmetic", 2: "UpAndOut"}
        extra = f", barrier={self.barrier}, rebate={self.rebate}" if self.kind == 2 else ""
        return f"Payoff.{names[self.kind]}(strike={self.strike}, opt_type={self.opt_type.name}{extra})"


@dataclasses.dataclass
class McConfig:
    """mc.rs:48-64; the defaults are the reference's McConfig::default()."""
    n_paths: int = 200_000
    n_steps: int = 252
    seed: int = 0xC0FFEE
    antithetic: bool = True
    scheme: VarianceScheme = VarianceScheme.FullTruncationEuler

    def _c(self):
        c = _Config()
        c.n_paths, c.n_steps, c.seed = int(self.n_paths), int(self.n_steps), int(self.seed)
        c.antithetic = 1 if self.antithetic else 0
        c.scheme = int(self.scheme)
        return c


# Immutable records; named tuples, since a surface or batch call builds one
# per price (328 of each for a 41 x 8 surface, 2.5x faster than frozen
# dataclasses).
class McResult(NamedTuple):
    """mc.rs:28-32"""
    price: float
    std_error: float


class McStats(NamedTuple):
    """(sum, sum_sq, count) of mc.rs:164-196 plus the simulated path-steps."""
    sum: float
    sum_sq: float
    count: int
    path_steps: int


class McOptions:
    """Extension options (precision, replay draws, per-path outputs)."""

    def __init__(self, precision=PRECISION_FP64, replay=None, want_paths=False, barrier_fp64=False):
        self.precision = precision
        self.replay = replay  # (za, zb, jump_sum) as [n_steps, n_pairs] arrays, jump_sum may be None
        self.want_paths = want_paths
        self.barrier_fp64 = barrier_fp64  # OPE_FLAG_BARRIER_FP64


class _OptionsHolder:
    """Builds the C options struct and keeps its buffers alive."""

    def __init__(self, opts: Optional[McOptions], n_results: int, n_pairs: int = 0):
        o = opts or McOptions()
        self.c = _Options()
        self.c.precision = o.precision
        self.c.rng = RNG_REPLAY if o.replay is not None else RNG_PHILOX
        self.c.flags = FLAG_BARRIER_FP64 if getattr(o, "barrier_fp64", False) else 0
        self.stats = (_Stats * n_results)()
        self.c.stats = C.cast(self.stats, C.POINTER(_Stats))
        self.ms = C.c_double(0.0)
        self.c.device_ms = C.pointer(self.ms)
        self.kms = C.c_double(0.0)
        self.c.kernel_ms = C.pointer(self.kms)
        self.keep = []
        self.path_payoffs = self.path_states = None
        if o.replay is not None:
            za, zb, js = o.replay
            za = np.ascontiguousarray(za, dtype=np.float64)
            zb = np.ascontiguousarray(zb, dtype=np.float64)
            self.keep += [za, zb]
            self.rp = _Replay()
            self.rp.za = za.ctypes.data_as(C.POINTER(C.c_double))
            self.rp.zb = zb.ctypes.data_as(C.POINTER(C.c_double))
            if js is not None:
                js = np.ascontiguousarray(js, dtype=np.float64)
                self.keep.append(js)
                self.rp.jump_sum = js.ctypes.data_as(C.POINTER(C.c_double))
            self.rp.n_steps, self.rp.n_pairs = za.shape
            self.c.replay = C.pointer(self.rp)
            if o.want_paths:
                self.path_payoffs = np.full((za.shape[1], 2), np.nan)
                self.path_states = np.full((za.shape[1], 2, 4), np.nan)
                self.c.path_payoffs = self.path_payoffs.ctypes.data_as(C.POINTER(C.c_double))
                self.c.path_states = self.path_states.ctypes.data_as(C.POINTER(C.c_double))

    def stats_list(self):
        v = np.frombuffer(self.stats, _STATS_NP)
        return [McStats(*s) for s in zip(v["sum"].tolist(), v["sum_sq"].tolist(), v["count"].tolist(),
                                         v["path_steps"].tolist())]


def mc_heston(spot, expiry, rate, div_yield, params: HestonParams, payoff: Payoff, cfg: McConfig) -> McResult:
    """mc.rs:98-103 on the GPU (FP64, Philox draws)."""
    out = _Result()
    _check(lib().ope_mc_heston(spot, expiry, rate, div_yield, C.byref(params._c()), C.byref(payoff._c()),
                               C.byref(cfg._c()), C.byref(out)))
    return McResult(out.price, out.std_error)


def mc_bates(spot, expiry, rate, div_yield, params: HestonParams, lambda_, mu_j, sigma_j, payoff: Payoff,
             cfg: McConfig) -> McResult:
    """mc.rs:107-113 on the GPU (FP64, Philox draws)."""
    out = _Result()
    _check(lib().ope_mc_bates(spot, expiry, rate, div_yield, C.byref(params._c()), lambda_, mu_j, sigma_j,
                              C.byref(payoff._c()), C.byref(cfg._c()), C.byref(out)))
    return McResult(out.price, out.std_error)


@dataclasses.dataclass
class McRun:
    """Result of an extended call: the McResult plus its accumulators, the GPU
    time, and (replay mode) per-path payoffs [n_pairs, 2] and states
    [n_pairs, 2, 4] = (ln S_T, v_T, running_sum, running_max)."""
    result: McResult
    stats: McStats
    device_ms: float
    kernel_ms: float
    path_payoffs: Optional[np.ndarray] = None
    path_states: Optional[np.ndarray] = None


def mc_price_ex(spot, expiry, rate, div_yield, params: HestonParams, payoff: Payoff, cfg: McConfig,
                jumps=None, options: Optional[McOptions] = None) -> McRun:
    """mc_heston (jumps None) or mc_bates (jumps = (lambda, mu_j, sigma_j)) with options."""
    h = _OptionsHolder(options, 1)
    out = _Result()
    L = lib()
    if jumps is None:
        rc = L.ope_mc_heston_ex(spot, expiry, rate, div_yield, C.byref(params._c()), C.byref(payoff._c()),
                                C.byref(cfg._c()), C.byref(h.c), C.byref(out))
    else:
        lam, mu_j, sigma_j = jumps
        rc = L.ope_mc_bates_ex(spot, expiry, rate, div_yield, C.byref(params._c()), lam, mu_j, sigma_j,
                               C.byref(payoff._c()), C.byref(cfg._c()), C.byref(h.c), C.byref(out))
    _check(rc)
    return McRun(McResult(out.price, out.std_error), h.stats_list()[0], h.ms.value, h.kms.value, h.path_payoffs,
                 h.path_states)


def mc_surface(spot, rate, div_yield, params: HestonParams, jumps, opt_type, expiries: Sequence[float],
               strikes: Sequence[float], cfg: McConfig, precision=PRECISION_FP64, with_kernel_ms=False):
    """European strike x expiry grid from shared paths.  Returns
    (prices [E, K], std_errors [E, K], stats list, device_ms), plus the
    path-kernel time when with_kernel_ms."""
    lam, mu_j, sigma_j = jumps if jumps is not None else (0.0, 0.0, 0.0)
    ex = np.ascontiguousarray(expiries, dtype=np.float64)
    ks = np.ascontiguousarray(strikes, dtype=np.float64)
    n = ex.size * ks.size
    h = _OptionsHolder(McOptions(precision), n)
    out = (_Result * n)()
    _check(lib().ope_mc_surface(spot, rate, div_yield, C.byref(params._c()), lam, mu_j, sigma_j, int(opt_type),
                                ex.ctypes.data_as(C.POINTER(C.c_double)), ex.size,
                                ks.ctypes.data_as(C.POINTER(C.c_double)), ks.size, C.byref(cfg._c()),
                                C.byref(h.c), out))
    v = np.frombuffer(out, _RESULT_NP)
    prices = v["price"].reshape(ex.size, ks.size).copy()
    ses = v["std_error"].reshape(ex.size, ks.size).copy()
    if with_kernel_ms:
        return prices, ses, h.stats_list(), h.ms.value, h.kms.value
    return prices, ses, h.stats_list(), h.ms.value


@dataclasses.dataclass
class McRequest:
    """One request of a batch: mc_bates's arguments (lambda_ == 0: mc_heston)."""
    spot: float
    expiry: float
    rate: float
    div_yield: float
    params: HestonParams
    payoff: Payoff
    cfg: McConfig
    lambda_: float = 0.0
    mu_j: float = 0.0
    sigma_j: float = 0.0

    def _c(self):
        r = _Request()
        r.spot, r.expiry, r.rate, r.div_yield = self.spot, self.expiry, self.rate, self.div_yield
        r.heston = self.params._c()
        r.lambda_, r.mu_j, r.sigma_j = self.lambda_, self.mu_j, self.sigma_j
        r.payoff = self.payoff._c()
        r.cfg = self.cfg._c()
        return r

    def _row(self):
        """The fields of ope_mc_request in its order (a row of _REQUEST_NP)."""
        p, q, c = self.params, self.payoff, self.cfg
        return (self.spot, self.expiry, self.rate, self.div_yield, p.v0, p.kappa, p.theta, p.sigma, p.rho,
                self.lambda_, self.mu_j, self.sigma_j, q.kind, int(q.opt_type), q.strike, q.barrier, q.rebate,
                c.n_paths, c.n_steps, c.seed, 1 if c.antithetic else 0, int(c.scheme))


def mc_batch(requests: Sequence[McRequest], precision=PRECISION_FP64, with_kernel_ms=False):
    """All requests in one kernel launch.  Returns (results, stats, device_ms),
    plus the path-kernel time when with_kernel_ms."""
    n = len(requests)
    arr = (_Request * max(n, 1))()   # zeroed; filled through a numpy view of its bytes
    np.frombuffer(arr, _REQUEST_NP)[:n] = [r._row() for r in requests]
    h = _OptionsHolder(McOptions(precision), max(n, 1))
    out = (_Result * max(n, 1))()
    _check(lib().ope_mc_batch(arr, n, C.byref(h.c), out))
    v = np.frombuffer(out, _RESULT_NP)[:n]
    res = [McResult(p, s) for p, s in zip(v["price"].tolist(), v["std_error"].tolist())]
    if with_kernel_ms:
        return res, h.stats_list()[:n], h.ms.value, h.kms.value
    return res, h.stats_list()[:n], h.ms.value


def device_info():
    info = _DeviceInfo()
    _check(lib().ope_get_device_info(C.byref(info)))
    return {"name": info.name.decode(), "arch": info.arch.decode(), "compute_units": info.compute_units,
            "clock_mhz": info.clock_mhz, "total_mem_bytes": info.total_mem_bytes}


def debug_philox(seed: int, ctr: np.ndarray) -> np.ndarray:
    """Philox4x32-10 of counters ctr [n, 4] (uint32) under key seed, on the GPU."""
    ctr = np.ascontiguousarray(ctr, dtype=np.uint32).reshape(-1, 4)
    out = np.zeros_like(ctr)
    _check(lib().ope_debug_philox(seed, ctr.ctypes.data_as(C.POINTER(C.c_uint32)), ctr.shape[0],
                                  out.ctypes.data_as(C.POINTER(C.c_uint32))))
    return out


def debug_fp64_math(fn: int, x) -> np.ndarray:
    """The FP64 path's elementary functions on the GPU: fn 0 log, 1 exp,
    2 / 3 sin / cos of 2 pi x 2^-23 (integer x), 4 reciprocal, 5 reciprocal
    square root, 6 square root, 7 normal tail Phi(-|x|) for |x| <= 6.8."""
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    out = np.zeros_like(x)
    P = C.POINTER(C.c_double)
    _check(lib().ope_debug_fp64_math(int(fn), x.ctypes.data_as(P), x.size, out.ctypes.data_as(P)))
    return out


def debug_draws(seed, pair0, n_pairs, n_steps, expiry=1.0, lambda_=0.0, mu_j=0.0, sigma_j=0.0,
                precision=PRECISION_FP64):
    """The Philox-mode draws the pricing kernels use, [n_steps, n_pairs] each."""
    za = np.zeros((n_steps, n_pairs))
    zb = np.zeros((n_steps, n_pairs))
    js = np.zeros((n_steps, n_pairs))
    nj = np.zeros((n_steps, n_pairs), dtype=np.uint32)
    P = C.POINTER(C.c_double)
    _check(lib().ope_debug_draws(seed, pair0, n_pairs, n_steps, expiry, lambda_, mu_j, sigma_j, precision,
                                 za.ctypes.data_as(P), zb.ctypes.data_as(P), js.ctypes.data_as(P),
                                 nj.ctypes.data_as(C.POINTER(C.c_uint32))))
    return za, zb, js, nj


# ------------------------------------------------------------------ serving

class _ServerConfig(C.Structure):
    _fields_ = [("precision", C.c_int32), ("max_requests", C.c_uint32), ("memory_budget_bytes", C.c_uint64),
                ("initial_round_pairs", C.c_uint32), ("ring_items", C.c_uint32), ("idle_exit_ms", C.c_double),
                ("watchdog_ms", C.c_double)]


class _ServerRequestInfo(C.Structure):
    _fields_ = [("paths", C.c_uint64), ("path_steps", C.c_uint64), ("rounds", C.c_uint32), ("slot", C.c_uint32),
                ("queue_ms", C.c_double), ("latency_ms", C.c_double), ("stop_reason", C.c_int32),
                ("pad_", C.c_uint32)]


class _ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("path_steps", C.c_uint64),
                ("peak_resident_bytes", C.c_uint64), ("peak_active", C.c_uint32), ("kernel_launches", C.c_uint32),
                ("workgroups", C.c_uint32), ("pad_", C.c_uint32)]


def _server_lib():
    L = lib()
    if not getattr(L, "_srv_ready", False):
        L.ope_server_create.argtypes = [C.POINTER(_ServerConfig), C.POINTER(C.c_void_p)]
        L.ope_server_submit.argtypes = [C.c_void_p, C.POINTER(_Request), C.c_double, C.POINTER(C.c_uint64)]
        L.ope_server_submit_ex.argtypes = [C.c_void_p, C.POINTER(_Request), C.c_double, C.c_double,
                                           C.POINTER(C.c_uint64)]
        L.ope_server_cancel.argtypes = [C.c_void_p, C.c_uint64]
        L.ope_server_poll.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(_Result), C.POINTER(_ServerRequestInfo),
                                      C.POINTER(C.c_int32)]
        L.ope_server_wait.argtypes = [C.c_void_p, C.c_uint64, C.c_double, C.POINTER(_Result),
                                      C.POINTER(_ServerRequestInfo)]
        L.ope_server_get_stats.argtypes = [C.c_void_p, C.POINTER(_ServerStats)]
        L.ope_server_destroy.argtypes = [C.c_void_p]
        for n in ("ope_server_create", "ope_server_submit", "ope_server_submit_ex", "ope_server_cancel",
                  "ope_server_poll", "ope_server_wait", "ope_server_get_stats", "ope_server_destroy"):
            getattr(L, n).restype = C.c_int
        L._srv_ready = True
    return L


@dataclasses.dataclass(frozen=True)
class ServedResult:
    result: McResult
    paths: int
    path_steps: int
    rounds: int
    slot: int
    queue_ms: float
    latency_ms: float
    stop_reason: int = 0   # SERVER_STOP_*


class Server:
    """Persistent-kernel request loop (ope_server_*).  submit() queues a
    request (optionally with a target relative standard error), wait()/poll()
    return its result; results do not depend on what else is being served."""

    def __init__(self, precision=PRECISION_FP32_PATH, max_requests=1024, memory_budget_bytes=64 << 30,
                 initial_round_pairs=65536, idle_exit_ms=200.0, ring_items=0, watchdog_ms=0.0):
        L = _server_lib()
        cfg = _ServerConfig(precision, max_requests, memory_budget_bytes, initial_round_pairs, ring_items,
                            idle_exit_ms, watchdog_ms)
        h = C.c_void_p()
        _check(L.ope_server_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self.precision = precision

    def submit(self, req: McRequest, target_rel_se: float = 0.0, deadline_ms: float = 0.0) -> int:
        """Queues a request; deadline_ms > 0 sets a deadline that many
        milliseconds from now (ope_server_submit_ex)."""
        t = C.c_uint64()
        r = req._c()
        if deadline_ms > 0.0:
            _check(_server_lib().ope_server_submit_ex(self._h, C.byref(r), target_rel_se, deadline_ms, C.byref(t)))
        else:
            _check(_server_lib().ope_server_submit(self._h, C.byref(r), target_rel_se, C.byref(t)))
        return t.value

    def cancel(self, ticket: int) -> None:
        _check(_server_lib().ope_server_cancel(self._h, ticket))

    @staticmethod
    def _served(res, info):
        return ServedResult(McResult(res.price, res.std_error), info.paths, info.path_steps, info.rounds, info.slot,
                            info.queue_ms, info.latency_ms, info.stop_reason)

    def poll(self, ticket: int):
        res, info, done = _Result(), _ServerRequestInfo(), C.c_int32()
        _check(_server_lib().ope_server_poll(self._h, ticket, C.byref(res), C.byref(info), C.byref(done)))
        return self._served(res, info) if done.value else None

    def wait(self, ticket: int, timeout_ms: float = -1.0) -> ServedResult:
        res, info = _Result(), _ServerRequestInfo()
        _check(_server_lib().ope_server_wait(self._h, ticket, timeout_ms, C.byref(res), C.byref(info)))
        return self._served(res, info)

    def stats(self) -> dict:
        st = _ServerStats()
        _check(_server_lib().ope_server_get_stats(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in _ServerStats._fields_ if k != "pad_"}

    def close(self):
        if self._h:
            _server_lib().ope_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------- analytic pricers
from .cf import (CF_DTYPE, CONTRACT_DTYPE, GREEKS_DTYPE, BatesParams, CfOption, OptionContract,  # noqa: E402
                 PricingResult, batch_bates, batch_bates_greeks, batch_bsm, batch_bsm_price, batch_heston,
                 batch_heston_greeks, batch_implied_vol, bates_price, bates_price_and_greeks, bsm_chain,
                 bsm_price, bsm_price_and_greeks, cf_price_options, heston_price, heston_price_and_greeks,
                 implied_vol, implied_vol_chain, CALIB_DTYPE, calib_jacobian, calib_residuals)
