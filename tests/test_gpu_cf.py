"""Analytic Heston / Bates pricer on the GPU (include/ope_cf.h) against the
oracle's C restatement of heston.rs / bates.rs (itself pinned against an
independent Lewis integral in test_oracle_prices.py), and the reference's own
analytic-pricer tests (heston.rs:296-320, bates.rs:95-170)."""
import itertools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PARAM_SETS = [
    dict(v0=0.04, kappa=2.0, theta=0.04, sigma=0.3, rho=-0.7),   # the reference's test set (Feller holds)
    dict(v0=0.04, kappa=1.5, theta=0.04, sigma=0.5, rho=-0.7),   # the MC workloads' set (Feller violated)
    dict(v0=0.09, kappa=1.0, theta=0.09, sigma=0.8, rho=-0.5),
    dict(v0=0.02, kappa=4.0, theta=0.06, sigma=0.2, rho=0.3),
]
JUMP_SETS = [(0.5, -0.1, 0.15), (1.0, -0.2, 0.25)]
STRIKES = [60.0, 80.0, 95.0, 100.0, 105.0, 120.0, 150.0]
EXPIRIES = [0.02, 0.1, 0.5, 1.0, 2.0, 5.0]

# The device and the host take the same bisection decisions; prices differ
# only through last-bit differences of exp / log / sin / cos / atan2 / hypot,
# which the quadrature sums over up to ~10^3 nodes (measured on MI355X: at
# most 5e-14 at S0 = 100, two thirds of the prices bit-identical).
TOL = 1e-12  # relative to the spot


def grid(ope, with_jumps):
    opts, refs = [], []
    for ps, k, t, q, ot in itertools.product(PARAM_SETS, STRIKES, EXPIRIES, (0.0, 0.02), (0, 1)):
        hp = ope.HestonParams(**ps)
        c = ope.OptionContract(100.0, k, t, 0.03, q, 0.0, ope.OptionType(ot))
        if with_jumps:
            for lam, mj, sj in JUMP_SETS:
                opts.append(ope.CfOption(c, ope.BatesParams(hp, lam, mj, sj)))
                refs.append(("bates", ps, k, t, q, ot, (lam, mj, sj)))
        else:
            opts.append(ope.CfOption(c, hp))
            refs.append(("heston", ps, k, t, q, ot, None))
    return opts, refs


def oracle_price(ora, ref):
    kind, ps, k, t, q, ot, jp = ref
    hp = ora.heston(**ps)
    if kind == "heston":
        return ora.heston_price(100.0, k, t, 0.03, q, hp, ot)
    return ora.bates_price(100.0, k, t, 0.03, q, hp, *jp, ot)


@pytest.mark.parametrize("with_jumps", [False, True])
def test_matches_oracle_over_grid(ope, ora, with_jumps):
    opts, refs = grid(ope, with_jumps)
    print(f"cf grid jumps={with_jumps} n={len(opts)}", flush=True)
    got, panels = ope.cf_price_options(opts, with_panels=True)
    want = np.array([oracle_price(ora, r) for r in refs])
    err = np.abs(got - want)
    worst = int(np.argmax(err))
    print(f"max |gpu - oracle| = {err.max():.3e} (median {np.median(err):.1e}); "
          f"{int(np.sum(err == 0))} of {len(err)} bit-identical; panels {panels.min()}..{panels.max()}", flush=True)
    assert err.max() <= TOL * 100.0, (refs[worst], got[worst], want[worst])
    assert panels.min() >= 1 and panels.max() <= 200


def test_single_and_batch_entry_points_agree(ope):
    hp = ope.HestonParams(0.04, 1.5, 0.04, 0.5, -0.7)
    bp = ope.BatesParams(hp, 0.5, -0.1, 0.15)
    chain = [ope.OptionContract(100.0, 60.0 + 2.0 * i, 0.5, 0.03, 0.0, 0.2,
                                ope.OptionType.Call if i % 2 else ope.OptionType.Put) for i in range(41)]
    bh = ope.batch_heston(chain, hp)
    bb = ope.batch_bates(chain, bp)
    for i, c in enumerate(chain):
        assert bh[i] == ope.heston_price(c.spot, c.strike, c.expiry, c.rate, c.div_yield, hp, c.opt_type)
        assert bb[i] == ope.bates_price(c.spot, c.strike, c.expiry, c.rate, c.div_yield, bp, c.opt_type)


def test_reference_put_call_parity_and_sanity(ope):
    # heston.rs:296-314
    p = ope.HestonParams(0.04, 2.0, 0.04, 0.3, -0.7)
    call = ope.heston_price(100.0, 100.0, 0.5, 0.03, 0.0, p, ope.OptionType.Call)
    put = ope.heston_price(100.0, 100.0, 0.5, 0.03, 0.0, p, ope.OptionType.Put)
    assert abs(call - put - (100.0 - 100.0 * np.exp(-0.03 * 0.5))) < 0.01
    call = ope.heston_price(100.0, 100.0, 1.0, 0.05, 0.0, p, ope.OptionType.Call)
    put = ope.heston_price(100.0, 100.0, 1.0, 0.05, 0.0, p, ope.OptionType.Put)
    assert abs(call - put - 100.0 + 100.0 * np.exp(-0.05)) < 0.05 and call >= 0.0
    # bates.rs:108-115
    bp = ope.BatesParams(p, 0.5, -0.1, 0.15)
    call = ope.bates_price(100.0, 100.0, 0.5, 0.03, 0.0, bp, ope.OptionType.Call)
    put = ope.bates_price(100.0, 100.0, 0.5, 0.03, 0.0, bp, ope.OptionType.Put)
    assert abs(call - put - (100.0 - 100.0 * np.exp(-0.03 * 0.5))) < 0.02


def test_reference_bates_recovers_heston_without_jumps(ope):
    # bates.rs:98-105: lambda = 0 (and a vanishing jump volatility)
    p = ope.HestonParams(0.04, 2.0, 0.04, 0.3, -0.7)
    b = ope.bates_price(100.0, 100.0, 1.0, 0.05, 0.0, ope.BatesParams(p, 0.0, 0.0, 1e-8))
    h = ope.heston_price(100.0, 100.0, 1.0, 0.05, 0.0, p)
    assert abs(b - h) < 1e-12


def test_reference_bates_no_static_arbitrage(ope):
    # bates.rs:140-168: intrinsic <= C <= S e^{-qT}, C non-increasing in K
    sets = [ope.BatesParams(ope.HestonParams(0.04, 2.0, 0.04, 0.3, -0.7), 0.5, -0.1, 0.15),
            ope.BatesParams(ope.HestonParams(0.09, 1.0, 0.09, 0.8, -0.5), 1.0, -0.2, 0.25)]
    s, r, q = 100.0, 0.03, 0.0
    expiries = [0.02, 0.1, 0.25, 0.5, 1.0, 2.0]
    strikes = [70.0, 85.0, 100.0, 115.0, 130.0, 150.0, 200.0]
    for bp in sets:
        chain = [ope.OptionContract(s, k, t, r, q) for t in expiries for k in strikes]
        px = ope.batch_bates(chain, bp).reshape(len(expiries), len(strikes))
        for i, t in enumerate(expiries):
            prev = np.inf
            for j, k in enumerate(strikes):
                c = px[i, j]
                intrinsic = max(s * np.exp(-q * t) - k * np.exp(-r * t), 0.0)
                assert c >= intrinsic - 1e-4 and c <= s * np.exp(-q * t) + 1e-4 and c <= prev + 1e-4
                prev = c


def test_zero_vol_of_vol_is_bsm(ope, ora):
    # heston.rs:407-426: sigma -> 0 with v0 = theta collapses onto BSM
    p = ope.HestonParams(0.04, 1.0, 0.04, 1e-4, 0.0)
    chain = [ope.OptionContract(100.0, k, t, 0.03, 0.0) for t in (0.25, 1.0, 2.0)
             for k in (70.0, 85.0, 100.0, 115.0, 130.0)]
    for c, got in zip(chain, ope.batch_heston(chain, p)):
        assert abs(got - ora.bsm_price(100.0, c.strike, c.expiry, 0.03, 0.0, 0.2)) < 2e-3


def test_mc_matches_analytic_on_the_gpu(ope):
    # the two GPU paths against each other: Bates Euler MC within 3 SE of the CF price
    p = ope.HestonParams(0.04, 2.0, 0.04, 0.3, -0.7)
    cfg = ope.McConfig(n_paths=2_000_000, n_steps=200, seed=7)
    mc = ope.mc_bates(100.0, 1.0, 0.03, 0.0, p, 0.5, -0.1, 0.15, ope.Payoff.European(100.0), cfg)
    cf = ope.bates_price(100.0, 100.0, 1.0, 0.03, 0.0, ope.BatesParams(p, 0.5, -0.1, 0.15))
    # Euler discretisation bias at 200 steps is well below 3 SE here
    assert abs(mc.price - cf) < 3.0 * mc.std_error + 0.02


def test_record_array_input_is_the_object_input(ope):
    hp = ope.HestonParams(0.05, 2.5, 0.03, 0.6, -0.4)
    objs = [ope.CfOption(ope.OptionContract(100.0, k, t, 0.03, 0.01, 0.0, ope.OptionType(ot)),
                         ope.BatesParams(hp, 0.8, -0.05, 0.2) if m else hp)
            for m in (0, 1) for t in (0.3, 1.7) for k in (90.0, 110.0) for ot in (0, 1)]
    recs = np.zeros(len(objs), dtype=ope.CF_DTYPE)
    for i, o in enumerate(objs):
        c, p = o.contract, o.params
        h = p.heston if isinstance(p, ope.BatesParams) else p
        recs[i] = (c.spot, c.strike, c.expiry, c.rate, c.div_yield, 0.0, int(c.opt_type), 0, h.v0, h.kappa, h.theta,
                   h.sigma, h.rho, *((p.lambda_, p.mu_j, p.sigma_j) if isinstance(p, ope.BatesParams) else (0, 0, 0)),
                   int(isinstance(p, ope.BatesParams)), 0)
    assert np.array_equal(ope.cf_price_options(recs), ope.cf_price_options(objs))


GREEKS = ("price", "delta", "gamma", "vega", "theta", "rho", "vanna", "volga")


@pytest.mark.parametrize("model", ["heston", "bates"])
def test_greeks_match_the_oracle(ope, ora, model):
    # greeks.rs:19-82 through the GPU pricer against the same bump-and-reprice
    # through the oracle's pricer: the 14 bumped prices agree to ~5e-14, the
    # difference quotients divide that by at most (2 dv)^2 = 4e-4
    ps = PARAM_SETS[1]
    hp, ohp = ope.HestonParams(**ps), ora.heston(**ps)
    chain = [ope.OptionContract(100.0, k, t, 0.03, q, 0.0, ope.OptionType(ot))
             for k in (80.0, 100.0, 125.0) for t in (0.1, 1.0, 3.0) for q in (0.0, 0.02) for ot in (0, 1)]
    if model == "heston":
        got = ope.batch_heston_greeks(chain, hp)
    else:
        got = ope.batch_bates_greeks(chain, ope.BatesParams(hp, 0.5, -0.1, 0.15))
    worst = 0.0
    for c, g in zip(chain, got):
        want = ora.cf_greeks(c.spot, c.strike, c.expiry, c.rate, c.div_yield, ohp,
                             None if model == "heston" else (0.5, -0.1, 0.15), int(c.opt_type))
        for name in GREEKS:
            err = abs(getattr(g, name) - want[name])
            worst = max(worst, err)
            assert err < 1e-8 * max(1.0, abs(want[name])), (c, name, getattr(g, name), want[name])
    print(f"greeks {model}: max |gpu - oracle| = {worst:.2e}", flush=True)


def test_greeks_single_calls_equal_the_batch(ope):
    hp = ope.HestonParams(0.04, 2.0, 0.04, 0.3, -0.7)
    bp = ope.BatesParams(hp, 0.5, -0.1, 0.15)
    chain = [ope.OptionContract(100.0, k, 1.0, 0.05, 0.0) for k in (90.0, 100.0, 110.0)]
    for c, g in zip(chain, ope.batch_heston_greeks(chain, hp)):
        assert g == ope.heston_price_and_greeks(c.spot, c.strike, c.expiry, c.rate, c.div_yield, hp)
    for c, g in zip(chain, ope.batch_bates_greeks(chain, bp)):
        assert g == ope.bates_price_and_greeks(c.spot, c.strike, c.expiry, c.rate, c.div_yield, bp)
        assert g.price == ope.bates_price(c.spot, c.strike, c.expiry, c.rate, c.div_yield, bp)


def test_reference_greeks_signs(ope):
    # heston.rs:322-339, bates.rs:118-136
    p = ope.HestonParams(0.04, 2.0, 0.04, 0.3, -0.7)
    bp = ope.BatesParams(p, 0.5, -0.1, 0.15)
    for r in (ope.heston_price_and_greeks(100.0, 100.0, 1.0, 0.05, 0.0, p),
              ope.bates_price_and_greeks(100.0, 100.0, 1.0, 0.05, 0.0, bp)):
        assert 0.0 < r.delta < 1.0 and r.gamma > 0.0 and r.vega > 0.0 and r.rho > 0.0
    for r in (ope.heston_price_and_greeks(100.0, 100.0, 1.0, 0.05, 0.0, p, ope.OptionType.Put),
              ope.bates_price_and_greeks(100.0, 100.0, 1.0, 0.05, 0.0, bp, ope.OptionType.Put)):
        assert -1.0 < r.delta < 0.0 and r.gamma > 0.0
    # bates.rs:140-148: without jumps the Greeks follow Heston's
    b = ope.bates_price_and_greeks(100.0, 100.0, 1.0, 0.05, 0.0, ope.BatesParams(p, 0.0, 0.0, 1e-8))
    h = ope.heston_price_and_greeks(100.0, 100.0, 1.0, 0.05, 0.0, p)
    assert abs(b.delta - h.delta) < 0.01 and abs(b.vega - h.vega) < 1.0


def _quotes(ope, ora, hp, jumps=None):
    rows = []
    for t in (0.25, 0.75, 1.5):
        for k in (85.0, 95.0, 100.0, 105.0, 120.0):
            ot = 0 if k >= 100.0 else 1
            h = ora.heston(*hp)
            px = (ora.bates_price(100.0, k, t, 0.03, 0.0, h, *jumps, ot) if jumps
                  else ora.heston_price(100.0, k, t, 0.03, 0.0, h, ot))
            rows.append(((100.0, k, t, 0.03, 0.0, 0.0, ot, 0), ora.implied_vol(100.0, k, t, 0.03, 0.0, ot, px),
                         1.0 + 0.1 * k / 100.0))
    return np.array(rows, dtype=ope.CALIB_DTYPE)


@pytest.mark.parametrize("model", ["heston", "bates"])
def test_calibration_evaluations_match_the_oracle(ope, ora, model):
    # calibration.rs:573-667 on the GPU against the oracle's restatement, at a
    # parameter vector away from the one that generated the quotes
    gen, jg = [0.04, 1.5, 0.05, 0.5, -0.7], [0.5, -0.1, 0.15]
    p0, j0 = [0.05, 2.0, 0.04, 0.4, -0.6], [0.4, -0.08, 0.12]
    q = _quotes(ope, ora, gen, jg if model == "bates" else None)
    hp = ope.HestonParams(*p0)
    params = ope.BatesParams(hp, *j0) if model == "bates" else hp
    res, jac, jtj, grad = ope.calib_jacobian(q, params)
    rows = np.column_stack([q["contract"][f] for f in ("spot", "strike", "expiry", "rate", "div_yield")]
                           + [q["contract"]["opt_type"], q["iv_market"], q["weight"]])
    want = ora.calib_eval(rows, 1 if model == "bates" else 0, p0 + (j0 if model == "bates" else []))
    for got, ref, name in zip((res, jac, jtj, grad), want, ("residuals", "jacobian", "jtj", "grad")):
        err = np.max(np.abs(got - ref) / np.maximum(1.0, np.abs(ref)))
        print(f"calibration {model} {name}: max rel |gpu - oracle| = {err:.2e}", flush=True)
        assert err < (1e-10 if name in ("residuals", "jacobian") else 1e-9), (name, got, ref)
    assert np.array_equal(ope.calib_residuals(q, params), res)
    if model == "bates":
        # the jump columns are AD's (calibration.rs:617-634): d price / d param
        # from the oracle's restatement of ad.rs:882, over the BSM vega at the
        # model's implied vol
        for i, row in enumerate(rows):
            iv = res[i] + row[6]
            vega = ora.bsm_greeks(row[0], row[1], row[2], row[3], row[4], iv, int(row[5]))["vega"]
            sens = ora.bates_jump_sens(row[0], row[1], row[2], row[3], row[4], ora.heston(*p0), *j0, int(row[5]))
            np.testing.assert_allclose(jac[i, 5:], sens[1:] / vega, rtol=1e-10, atol=1e-12)


def _surface_quotes(ope, model):
    # the calibration workload of the bench: 41 strikes 60..140 x 8 expiries
    # 0.1..2 priced at the generating parameters, their implied vols the quotes
    strikes = [60.0 + 2.0 * i for i in range(41)]
    expiries = [0.1, 0.25, 0.5, 0.75, 1.0, 1.25, 1.5, 2.0]
    chain = np.zeros(len(expiries) * len(strikes), dtype=ope.CONTRACT_DTYPE)
    chain["spot"] = 100.0
    chain["strike"] = np.tile(strikes, len(expiries))
    chain["expiry"] = np.repeat(expiries, len(strikes))
    chain["rate"] = 0.03
    opts = np.zeros(len(chain), dtype=ope.CF_DTYPE)
    for f in ("spot", "strike", "expiry", "rate", "div_yield", "opt_type"):
        opts[f] = chain[f]
    for f, v in zip(("v0", "kappa", "theta", "sigma", "rho"), (0.04, 1.5, 0.04, 0.5, -0.7)):
        opts[f] = v
    if model == "bates":
        opts["lambda_"], opts["mu_j"], opts["sigma_j"], opts["model"] = 0.5, -0.1, 0.15, 1
    px = ope.cf_price_options(opts)
    iv, found, _ = ope.implied_vol_chain(chain, px)
    q = np.zeros(len(chain), dtype=ope.CALIB_DTYPE)
    q["contract"], q["iv_market"], q["weight"] = chain, np.where(found == 1, iv, 0.2), 1.0
    return q


@pytest.mark.parametrize("model", ["heston", "bates"])
def test_calibration_on_the_41_by_8_surface(ope, ora, model):
    # 328 quotes including 0.1-year wings whose BSM vega is 1e-6..1e-4: there
    # the implied vol is fixed only to the solver's price tolerance, |BSM(iv)
    # - price| < 1e-10 (iv.rs:12, 69), so two correct solvers may return vols
    # 2e-10 / vega apart, and a differenced column (step h) may differ by
    # 4e-10 / (vega h).  Elsewhere GPU and oracle agree to 1e-8, and the AD
    # jump columns, which divide by vega instead of differencing vols, to 1e-10.
    q = _surface_quotes(ope, model)
    bates = model == "bates"
    p0, j0 = [0.05, 2.0, 0.045, 0.4, -0.6], [0.4, -0.08, 0.12]
    params = ope.BatesParams(ope.HestonParams(*p0), *j0) if bates else ope.HestonParams(*p0)
    res, jac, _, _ = ope.calib_jacobian(q, params)
    rows = np.column_stack([q["contract"][f] for f in ("spot", "strike", "expiry", "rate", "div_yield")]
                           + [q["contract"]["opt_type"], q["iv_market"], q["weight"]])
    pv = p0 + (j0 if bates else [])
    ref_res, ref_jac, _, _ = ora.calib_eval(rows, 1 if bates else 0, pv)
    vega = np.array([ora.bsm_greeks(r[0], r[1], r[2], r[3], r[4], rr + r[6], int(r[5]))["vega"]
                     for r, rr in zip(rows, ref_res)])
    bumps = np.array([1e-4, 1e-3, 1e-4, 1e-3, 1e-4, 1e-4, 1e-4, 1e-4][:len(pv)])
    slack_res = 2e-10 / vega
    assert np.all(np.abs(res - ref_res) <= 1e-12 + slack_res)
    slack = 4e-10 / (vega[:, None] * 2 * bumps[None, :])
    rel = np.abs(jac - ref_jac) / np.maximum(1.0, np.abs(ref_jac))
    assert np.all(np.abs(jac - ref_jac) <= 1e-8 * np.maximum(1.0, np.abs(ref_jac)) + slack)
    ok = vega >= 1e-2
    print(f"{model}: {ok.sum()} of {len(q)} quotes with vega >= 1e-2, max rel diff there {rel[ok].max():.2e}; "
          f"overall {rel.max():.2e} at vega {vega[np.argmax(rel.max(axis=1))]:.1e}", flush=True)
    assert rel[ok].max() < 1e-8
    if bates:
        assert rel[:, 5:].max() < 1e-10
