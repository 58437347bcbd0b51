"""The C ABI is blocking and thread-safe: calls made from several host
threads at once (ctypes releases the GIL for the duration of a foreign call)
return the same bits as the same calls made one after another, whichever
entry points they mix (Monte Carlo prices in both precisions, surfaces,
batches, analytic prices, BSM chains, implied volatilities)."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HP = dict(v0=0.04, kappa=1.5, theta=0.04, sigma=0.5, rho=-0.7)


def jobs(ope):
    hp = ope.HestonParams(**HP)
    bp = ope.BatesParams(hp, 0.5, -0.1, 0.15)
    chain = [ope.OptionContract(100.0, 60.0 + 2.0 * i, 0.5, 0.03, 0.0, 0.2,
                                ope.OptionType.Call if i % 2 else ope.OptionType.Put) for i in range(41)]
    contracts = np.zeros(4096, dtype=ope.CONTRACT_DTYPE)
    rng = np.random.default_rng(7)
    contracts["spot"] = 100.0
    contracts["strike"] = rng.uniform(60.0, 160.0, len(contracts))
    contracts["expiry"] = rng.uniform(0.05, 3.0, len(contracts))
    contracts["rate"] = 0.03
    contracts["vol"] = rng.uniform(0.05, 0.8, len(contracts))
    contracts["opt_type"] = rng.integers(0, 2, len(contracts))

    def mc(precision, scheme, jumps, payoff, seed):
        cfg = ope.McConfig(n_paths=200_000, n_steps=64, seed=seed, scheme=scheme)
        r = ope.mc_price_ex(100.0, 1.0, 0.03, 0.0, hp, payoff, cfg, jumps, ope.McOptions(precision)).result
        return (r.price, r.std_error)

    def surface(precision):
        cfg = ope.McConfig(n_paths=100_000, n_steps=32, seed=5)
        p, s, _, _ = ope.mc_surface(100.0, 0.03, 0.0, hp, (0.5, -0.1, 0.15), 0, [0.25, 1.0],
                                    np.arange(80.0, 121.0, 5.0), cfg, precision)
        return tuple(p.ravel()) + tuple(s.ravel())

    def batch(precision):
        reqs = [ope.McRequest(100.0, 1.0, 0.03, 0.0, hp, pay, ope.McConfig(n_paths=50_000 + 2 * i, n_steps=20 + i,
                                                                          seed=100 + i), lam, -0.1, 0.15)
                for i, (pay, lam) in enumerate([(ope.Payoff.European(100.0), 0.0),
                                                (ope.Payoff.AsianArithmetic(95.0), 0.5),
                                                (ope.Payoff.UpAndOut(100.0, 130.0), 0.0)])]
        res, _, _ = ope.mc_batch(reqs, precision)
        return tuple((r.price, r.std_error) for r in res)

    def bsm():
        g, _ = ope.bsm_chain(contracts, greeks=True)
        return g.tobytes()

    def iv():
        px, _ = ope.bsm_chain(contracts, greeks=False)
        vol, found, _ = ope.implied_vol_chain(contracts, px)
        return vol.tobytes() + found.tobytes()

    return [
        ("mc fp64 euler", lambda: mc(0, 0, None, ope.Payoff.European(100.0), 1)),
        ("mc fp32 qe asian", lambda: mc(1, 1, None, ope.Payoff.AsianArithmetic(100.0), 2)),
        ("mc fp32 bates uao", lambda: mc(1, 0, (0.5, -0.1, 0.15), ope.Payoff.UpAndOut(100.0, 130.0), 3)),
        ("mc fp64 qe bates", lambda: mc(0, 1, (0.5, -0.1, 0.15), ope.Payoff.European(90.0), 4)),
        ("surface fp32", lambda: surface(1)),
        ("batch fp64", lambda: batch(0)),
        ("batch fp32", lambda: batch(1)),
        ("cf heston chain", lambda: tuple(ope.batch_heston(chain, hp))),
        ("cf bates chain", lambda: tuple(ope.batch_bates(chain, bp))),
        ("bsm chain", bsm),
        ("implied vol chain", iv),
    ]


def test_concurrent_calls_equal_serial_calls(ope):
    work = jobs(ope)
    serial = {name: fn() for name, fn in work}
    print(f"serial results for {len(serial)} kinds of call", flush=True)
    n_threads, rounds = 6, 4
    got, errors = [], []
    lock = threading.Lock()

    def worker(t):
        try:
            for r in range(rounds):
                for k in range(len(work)):
                    name, fn = work[(k + t + r) % len(work)]   # threads interleave different entry points
                    out = fn()
                    with lock:
                        got.append((name, out))
        except Exception as e:  # reported below, after every thread has ended
            with lock:
                errors.append(repr(e))

    threads = [threading.Thread(target=worker, args=(t,)) for t in range(n_threads)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=100)
    assert not any(th.is_alive() for th in threads), "a host thread did not finish"
    assert not errors, errors
    assert len(got) == n_threads * rounds * len(work)
    bad = [name for name, out in got if out != serial[name]]
    print(f"{len(got)} concurrent calls, {len(bad)} differing from the serial results", flush=True)
    assert not bad, sorted(set(bad))
