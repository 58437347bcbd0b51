"""The C-ABI library loads, exports every symbol include/ope_mc.h declares,
has the struct layout the header fixes, and validates arguments on the host
(no GPU needed: invalid input is rejected before the device is touched)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "ope_mc.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ope_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(ope):
    L = ope.lib()
    names = declared_functions()
    assert "ope_mc_heston" in names and "ope_mc_bates" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_struct_layout(ope):
    assert C.sizeof(ope._Heston) == 40
    assert C.sizeof(ope._Payoff) == 32
    assert C.sizeof(ope._Config) == 32
    assert ope._Config.scheme.offset == 28
    assert C.sizeof(ope._Result) == 16
    assert C.sizeof(ope._Stats) == 32
    assert C.sizeof(ope._Request) == 4 * 8 + 40 + 3 * 8 + 32 + 32


def test_config_default_is_reference_default(ope):
    c = ope._Config()
    ope.lib().ope_mc_config_default(C.byref(c))
    d = ope.McConfig()
    assert (c.n_paths, c.n_steps, c.seed, c.antithetic, c.scheme) == (200000, 252, 0xC0FFEE, 1, 0)
    assert (d.n_paths, d.n_steps, d.seed, d.antithetic, int(d.scheme)) == (200000, 252, 0xC0FFEE, True, 0)


def _hp(ope, **kw):
    base = dict(v0=0.04, kappa=1.5, theta=0.04, sigma=0.5, rho=-0.7)
    base.update(kw)
    return ope.HestonParams(**base)


@pytest.mark.parametrize("case", [
    dict(cfg=dict(n_paths=1)),                       # antithetic needs >= 2 paths (mc.rs:160-161 gives NaN)
    dict(cfg=dict(n_paths=0, antithetic=False)),
    dict(cfg=dict(n_steps=0)),
    dict(cfg=dict(scheme=7)),
    dict(spot=-1.0),
    dict(spot=float("nan")),
    dict(expiry=-0.5),
    dict(hp=dict(sigma=0.0), cfg=dict(scheme=1)),      # QE divides by sigma (mc.rs:233-235)
    dict(hp=dict(kappa=0.0), cfg=dict(scheme=1)),      # and by kappa (mc.rs:293-294)
    dict(payoff=(5, 0)),                               # unknown payoff kind
    dict(payoff=(0, 3)),                               # unknown option type
    dict(jumps=(-0.1, 0.0, 0.1)),
    dict(jumps=(1e6, 0.0, 0.1)),                       # Poisson mean above the sampler limit
])
def test_invalid_arguments_rejected_on_host(ope, case):
    cfg = ope.McConfig(n_paths=1000, n_steps=10)
    for k, v in case.get("cfg", {}).items():
        setattr(cfg, k, v)
    kind, opt = case.get("payoff", (0, 0))
    pay = ope._Payoff(kind, opt, 100.0, 0.0, 0.0)
    hp = _hp(ope, **case.get("hp", {}))
    spot = case.get("spot", 100.0)
    expiry = case.get("expiry", 1.0)
    L = ope.lib()
    out = ope._Result()
    if "jumps" in case:
        lam, mu, sj = case["jumps"]
        rc = L.ope_mc_bates(spot, expiry, 0.03, 0.0, C.byref(hp._c()), lam, mu, sj, C.byref(pay), C.byref(cfg._c()),
                            C.byref(out))
    else:
        rc = L.ope_mc_heston(spot, expiry, 0.03, 0.0, C.byref(hp._c()), C.byref(pay), C.byref(cfg._c()), C.byref(out))
    assert rc == ope.OPE_ERR_INVALID_ARGUMENT
    assert L.ope_last_error_message()


def test_status_strings(ope):
    L = ope.lib()
    for code in range(0, -8, -1):
        assert L.ope_status_string(code)
    with pytest.raises(ope.OpeError) as e:
        ope.mc_heston(100.0, 1.0, 0.03, 0.0, _hp(ope), ope.Payoff.European(100.0), ope.McConfig(n_paths=1))
    assert e.value.status == ope.OPE_ERR_INVALID_ARGUMENT


def test_no_gpu_is_an_error_not_a_fallback(ope):
    import torch

    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(ope.OpeError) as e:
        ope.mc_heston(100.0, 1.0, 0.03, 0.0, _hp(ope), ope.Payoff.European(100.0), ope.McConfig(n_paths=100))
    assert e.value.status == ope.OPE_ERR_NO_DEVICE
