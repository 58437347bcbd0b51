"""ope_mc_bates_noalloc, the entry point that allocates nothing (its device
buffers live in a fixed array in the library's device code), called through
the C ABI: its results equal ope_mc_bates_ex's bit for bit, on consecutive
calls whose second is smaller than the first (the array keeps the previous
call's contents), and requests beyond its fixed maximum are rejected on the
host."""
import ctypes as C

import pytest

pytestmark = pytest.mark.gpu

HP = dict(v0=0.04, kappa=1.5, theta=0.04, sigma=0.5, rho=-0.7)
JUMPS = (0.5, -0.1, 0.15)


def _noalloc(ope, spot, expiry, rate, q, hp, jumps, payoff, cfg, precision):
    L = ope.lib()
    f = L.ope_mc_bates_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_double] * 4 + [C.POINTER(ope._Heston)] + [C.c_double] * 3 + \
                 [C.POINTER(ope._Payoff), C.POINTER(ope._Config), C.POINTER(ope._Options), C.POINTER(ope._Result)]
    h = ope._OptionsHolder(ope.McOptions(precision), 1)
    out = ope._Result()
    rc = f(spot, expiry, rate, q, C.byref(hp._c()), *jumps, C.byref(payoff._c()), C.byref(cfg._c()),
           C.byref(h.c), C.byref(out))
    return rc, ope.McResult(out.price, out.std_error), h.stats_list()[0]


CASES = [
    # (payoff, n_paths, n_steps, scheme, jumps): a large call (2M pairs, 4 pairs
    # per thread, 1954 tiles) followed by smaller ones of other configurations
    ("european", 4_000_000, 64, 0, JUMPS),
    ("asian", 100_000, 252, 1, (0.0, 0.0, 0.0)),
    ("up_and_out", 30_002, 90, 0, JUMPS),
]


def _payoff(ope, kind):
    return {"european": ope.Payoff.European(100.0), "asian": ope.Payoff.AsianArithmetic(95.0),
            "up_and_out": ope.Payoff.UpAndOut(100.0, 130.0, 1.5)}[kind]


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_equals_the_allocating_entry_point(ope, precision):
    hp = ope.HestonParams(**HP)
    for kind, n_paths, n_steps, scheme, jumps in CASES:
        print(f"no-alloc {kind} {n_paths} x {n_steps} scheme {scheme} precision {precision}", flush=True)
        cfg = ope.McConfig(n_paths=n_paths, n_steps=n_steps, seed=321, scheme=scheme)
        pay = _payoff(ope, kind)
        rc, got, st = _noalloc(ope, 100.0, 1.0, 0.03, 0.01, hp, jumps, pay, cfg, precision)
        assert rc == ope.OPE_OK, ope.lib().ope_last_error_message()
        ref = ope.mc_price_ex(100.0, 1.0, 0.03, 0.01, hp, pay, cfg, jumps, ope.McOptions(precision))
        assert got == ref.result
        assert st == ref.stats


def test_noalloc_rejects_requests_beyond_its_scratch(ope):
    hp = ope.HestonParams(**HP)
    # 2^24 + 1 pairs need 16385 tiles of 1024 pairs; the array holds 16384
    cfg = ope.McConfig(n_paths=2 * (2**24 + 1), n_steps=10, seed=1)
    rc, _, _ = _noalloc(ope, 100.0, 1.0, 0.03, 0.0, hp, JUMPS, ope.Payoff.European(100.0), cfg, 1)
    assert rc == ope.OPE_ERR_INVALID_ARGUMENT
    cfg = ope.McConfig(n_paths=2 * 2**24, n_steps=2, seed=1)   # exactly at the limit: runs
    rc, got, _ = _noalloc(ope, 100.0, 1.0, 0.03, 0.0, hp, JUMPS, ope.Payoff.European(100.0), cfg, 1)
    assert rc == ope.OPE_OK
    ref = ope.mc_price_ex(100.0, 1.0, 0.03, 0.0, hp, ope.Payoff.European(100.0), cfg, JUMPS, ope.McOptions(1))
    assert got == ref.result
