#!/usr/bin/env python3
"""Throughput of the Heston/Bates Monte Carlo path on one MI355X.

    python bench.py [--steps K] [--warmup W]
                    [--workload qe_asian_uao|bates_surface|heston_euler|uneven_batch|serving]

One step is one pass of the hot path over one batch of synthetic input:

* qe_asian_uao (default): Heston, Andersen QE (gamma1 = gamma2 = 1/2), S0=K=100,
  T=1, r=0.03, q=0, v0=theta=0.04, kappa=1.5, sigma=0.5, rho=-0.7; 1M paths
  (500k antithetic pairs) x 365 steps, seed 42; an arithmetic Asian call and an
  up-and-out call (B=130, daily monitoring), each its own simulation, in one
  batched call.
* bates_surface: the same Heston parameters plus lambda=0.5, mu_j=-0.1,
  sigma_j=0.15; 4M paths x 252 steps per expiry, 8 expiries, 41 call strikes
  60..140 priced from shared paths.
* heston_euler: the reference's CPU-sized case, 100k paths x 252 Euler steps.
* uneven_batch: 256 independent Heston/Bates requests of uneven path and step
  counts and mixed payoffs in one launch.
* serving: the persistent-kernel server under Poisson arrivals at 2000 req/s,
  each request run to a standard error of 1e-3 x price; p50/p99 latency.
* analytic_cf: the GPU characteristic-function pricer on 32 parameter sets x
  the 41 x 8 surface, Heston and Bates (20992 options per call), and the BSM
  price-and-Greeks and implied-volatility chains (4M and 1M options).

The default run also measures bates_surface, uneven_batch, analytic_cf,
bsm_iv and serving and carries them in the same JSON line.

The value is FP32-path mode path-steps/s (both antithetic siblings count);
FP64 mode runs the same workload on identical Philox streams and the line
carries its throughput and the FP32 price error in standard errors.  Kernel
times come from HIP events the library records on its own stream around the
path-simulation kernels.  The CPU baseline is the C restatement of the
reference engine (oracle/) timed on this host on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "options-pricing-engine-rs_amd"))

METRIC = "MC path-steps/s FP32 vs FP64 at matched stderr; p50/p99 request latency"
HP = dict(v0=0.04, kappa=1.5, theta=0.04, sigma=0.5, rho=-0.7)
JUMPS = (0.5, -0.1, 0.15)
EXPIRIES = [0.1, 0.25, 0.5, 0.75, 1.0, 1.25, 1.5, 2.0]
STRIKES = [60.0 + 2.0 * i for i in range(41)]


def make_workload(ope, wm, name):
    """Returns (description, config dict, run(precision) -> (prices, ses, kernel_ms), path_steps, flops)."""
    hp = ope.HestonParams(**HP)
    if name == "qe_asian_uao":
        cfg = ope.McConfig(n_paths=1_000_000, n_steps=365, seed=42, scheme=ope.VarianceScheme.QuadraticExponential)
        reqs = [ope.McRequest(100.0, 1.0, 0.03, 0.0, hp, ope.Payoff.AsianArithmetic(100.0), cfg),
                ope.McRequest(100.0, 1.0, 0.03, 0.0, hp, ope.Payoff.UpAndOut(100.0, 130.0, 0.0), cfg)]

        def run(precision):
            res, stats, dev_ms, k_ms = ope.mc_batch(reqs, precision, with_kernel_ms=True)
            return [r.price for r in res], [r.std_error for r in res], k_ms

        path_steps = 2 * 1_000_000 * 365
        flops = (wm.flops(1_000_000, 365, "qe", False, 1, "asian") + wm.flops(1_000_000, 365, "qe", False, 1, "up_and_out"))
        desc = ("Heston QE (gamma1=gamma2=0.5) 1M paths x 365 steps, arithmetic Asian call K=100 and "
                "up-and-out call K=100 B=130 (one simulation each), seed 42")
        conf = dict(n_paths=1_000_000, n_steps=365, scheme="QE", payoffs=["asian_call_100", "uao_call_100_130"],
                    simulations_per_step=2)
        return desc, conf, run, path_steps, flops
    if name == "bates_surface":
        cfg = ope.McConfig(n_paths=4_000_000, n_steps=252, seed=42)

        def run(precision):
            prices, ses, stats, dev_ms, k_ms = ope.mc_surface(100.0, 0.03, 0.0, hp, JUMPS, 0, EXPIRIES, STRIKES,
                                                              cfg, precision, with_kernel_ms=True)
            return list(prices.ravel()), list(ses.ravel()), k_ms

        path_steps = len(EXPIRIES) * 4_000_000 * 252
        flops = len(EXPIRIES) * wm.flops(4_000_000, 252, "euler", True, len(STRIKES), "european")
        desc = ("Bates (lambda=0.5, mu_j=-0.1, sigma_j=0.15) Euler 4M paths x 252 steps per expiry, "
                "8 expiries x 41 call strikes 60..140 from shared paths, seed 42")
        conf = dict(n_paths=4_000_000, n_steps=252, scheme="Euler", expiries=EXPIRIES, n_strikes=len(STRIKES))
        return desc, conf, run, path_steps, flops
    if name == "heston_euler":
        cfg = ope.McConfig(n_paths=100_000, n_steps=252, seed=42)

        def run(precision):
            r = ope.mc_price_ex(100.0, 1.0, 0.03, 0.0, hp, ope.Payoff.European(100.0), cfg, None,
                                ope.McOptions(precision))
            return [r.result.price], [r.result.std_error], r.kernel_ms

        desc = "Heston European call, full-truncation Euler, 100k paths x 252 steps, seed 42"
        return (desc, dict(n_paths=100_000, n_steps=252, scheme="Euler"), run, 100_000 * 252,
                wm.flops(100_000, 252, "euler", False))
    if name == "uneven_batch":
        reqs = draw_uneven_requests(ope, 256, 4)

        def run(precision):
            res, stats, dev_ms, k_ms = ope.mc_batch(reqs, precision, with_kernel_ms=True)
            return [r.price for r in res], [r.std_error for r in res], k_ms

        kinds = {0: "european", 1: "asian", 2: "up_and_out"}
        path_steps = sum(q.cfg.n_paths * q.cfg.n_steps for q in reqs)
        flops = sum(wm.flops(q.cfg.n_paths, q.cfg.n_steps, "euler", q.lambda_ > 0, 1, kinds[q.payoff.kind])
                    for q in reqs)
        desc = ("256 independent Heston/Bates requests in one launch, drawn as v0, theta ~ U[0.01, 0.16], "
                "kappa ~ U[0.5, 5], xi ~ U[0.1, 1], rho ~ U[-0.95, 0], lambda ~ U[0, 1], payoff uniform over "
                "{European, Asian, up-and-out}, paths log-uniform in [10k, 2M], steps uniform in [50, 500], Euler")
        conf = dict(n_requests=256, total_paths=sum(q.cfg.n_paths for q in reqs), scheme="Euler", draw_seed=4)
        return desc, conf, run, path_steps, flops
    raise SystemExit(f"unknown workload {name}")


def draw_uneven_requests(ope, n, seed):
    """Requests drawn as the uneven-batch configuration: v0, theta ~ U[0.01, 0.16],
    kappa ~ U[0.5, 5], xi ~ U[0.1, 1], rho ~ U[-0.95, 0], lambda ~ U[0, 1]
    (mu_j = -0.1, sigma_j = 0.15), payoff uniform over {European, Asian,
    up-and-out B=130}, K=100, S0=100, T=1, r=0.03, Euler, paths log-uniform in
    [1e4, 2e6] (even), steps uniform in [50, 500]."""
    import numpy as np

    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        v0, theta = rng.uniform(0.01, 0.16, 2)
        hp = ope.HestonParams(v0, rng.uniform(0.5, 5.0), theta, rng.uniform(0.1, 1.0), rng.uniform(-0.95, 0.0))
        kind = int(rng.integers(0, 3))
        pay = [ope.Payoff.European(100.0), ope.Payoff.AsianArithmetic(100.0),
               ope.Payoff.UpAndOut(100.0, 130.0, 0.0)][kind]
        paths = max(2, int(np.exp(rng.uniform(np.log(1e4), np.log(2e6)))) // 2 * 2)
        cfg = ope.McConfig(n_paths=paths, n_steps=int(rng.integers(50, 501)), seed=1000 + i)
        reqs.append(ope.McRequest(100.0, 1.0, 0.03, 0.0, hp, pay, cfg, float(rng.uniform(0.0, 1.0)), -0.1, 0.15))
    return reqs


def poisson_arrivals(rate, seconds, seed):
    """Arrival times (s) of a Poisson process at `rate` req/s over `seconds`;
    rate = inf: 2000 * seconds arrivals at t = 0."""
    import numpy as np

    rng = np.random.default_rng(seed)
    if math.isinf(rate):
        return np.zeros(int(2000.0 * seconds))
    gaps = rng.exponential(1.0 / rate, size=int(rate * seconds * 1.5) + 16)
    arrivals = np.cumsum(gaps)
    return arrivals[arrivals < seconds]


def submit_on_schedule(srv, arrivals, reqs, target):
    """Submits reqs[i] at arrivals[i] seconds after the start (host clock);
    returns (tickets, start time)."""
    tickets = []
    t0 = time.perf_counter()
    for a, q in zip(arrivals, reqs):
        while True:
            dt = t0 + a - time.perf_counter()
            if dt <= 0:
                break
            time.sleep(dt if dt > 2e-4 else 0)
        tickets.append(srv.submit(q, target))
    return tickets, t0


def serve(ope, precision, rate, seconds, target, seed=5):
    """Poisson arrivals at `rate` req/s for `seconds`; each request runs until its
    standard error <= target * price (cap: its drawn path count).  rate = inf:
    2000 * seconds requests arrive at once, which measures the server's
    saturated throughput."""
    import numpy as np

    arrivals = poisson_arrivals(rate, seconds, seed)
    reqs = draw_uneven_requests(ope, len(arrivals), seed + 1)
    with ope.Server(precision, max_requests=4096, memory_budget_bytes=64 << 30, idle_exit_ms=200.0) as srv:
        warm = srv.submit(reqs[0], target)   # first launch outside the measured window
        srv.wait(warm)
        tickets, t0 = submit_on_schedule(srv, arrivals, reqs, target)
        got = [srv.wait(t) for t in tickets]
        wall = time.perf_counter() - t0
        st = srv.stats()
    lat = np.array([g.latency_ms for g in got])
    ps = sum(g.path_steps for g in got)
    return {"requests": len(got), "offered_rate_per_s": rate if not math.isinf(rate) else "burst", "arrival_window_s": seconds,
            "target_rel_se": target, "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "max_ms": float(lat.max()),
            "path_steps_per_s": ps / wall, "wall_s": wall,
            "mean_rounds": float(np.mean([g.rounds for g in got])),
            "requests_by_rounds": {str(k): int(v) for k, v in zip(*np.unique([g.rounds for g in got],
                                                                             return_counts=True))},
            "stopped_at_target": int(sum(1 for g in got if g.stop_reason == ope.SERVER_STOP_TARGET)),
            "paths_over_cap": float(sum(g.paths for g in got) / sum(q.cfg.n_paths for q in reqs[:len(got)])),
            "met_target": int(sum(1 for g in got if g.result.std_error <= target * abs(g.result.price))),
            "peak_active": st["peak_active"], "peak_resident_bytes": st["peak_resident_bytes"],
            "kernel_launches": st["kernel_launches"], "workgroups": st["workgroups"]}


def secondary(ope, wm, name):
    """A second workload timed in the same run: FP32-path and FP64 throughput and error."""
    d, c, run, ps, fl = make_workload(ope, wm, name)
    # the GPU clock ramps over the first few calls of a workload (the Bates
    # surface: 5.6, 5.4, 5.2, 5.0, then 4.94-4.97 ms), so a few warm-up calls
    # precede the timed ones
    q32, s32, m32, k32 = time_mode(run, 1, 5, 4)
    q64, s64, m64, k64 = time_mode(run, 0, 3, 2)
    err = [abs(a - b) / s for a, b, s in zip(q32, q64, s64) if s > 0]
    a32 = fl / (k32 * 1e-3) / 1e12
    a64 = fl / (k64 * 1e-3) / 1e12
    print(f"# {name}: fp32 {m32:.2f} ms fp64 {m64:.2f} ms", file=sys.stderr, flush=True)
    return {"workload": d, "config": c, "path_steps": ps,
            "fp32_path_steps_per_s": ps / (m32 * 1e-3), "fp64_path_steps_per_s": ps / (m64 * 1e-3),
            "fp32_ms": m32, "fp64_ms": m64, "fp32_kernel_ms": k32, "fp64_kernel_ms": k64,
            "fp32_over_fp64": m64 / m32, "fp32_price_err_max_se": max(err),
            "roofline_fp32": {"bound": "valu", "achieved": a32, "peak": wm.PEAK_TFLOPS["f32"], "unit": "TFLOP/s",
                              "frac": a32 / wm.PEAK_TFLOPS["f32"]},
            "roofline_fp64": {"bound": "valu", "achieved": a64, "peak": wm.PEAK_TFLOPS["f64"], "unit": "TFLOP/s",
                              "frac": a64 / wm.PEAK_TFLOPS["f64"]}}


def cf_options(ope, seed=11, n_sets=32):
    """The analytic-pricer workload as an array of ope_cf_option records:
    n_sets parameter sets, each priced on the 41-strike x 8-expiry surface
    as Heston and as Bates: what a calibration sweep evaluates."""
    import numpy as np

    rng = np.random.default_rng(seed)
    rows = []
    for _ in range(n_sets):
        v0, theta = rng.uniform(0.02, 0.09, 2)
        kappa, sigma, rho = rng.uniform(0.5, 4.0), rng.uniform(0.2, 1.0), rng.uniform(-0.9, -0.1)
        for model in (0, 1):
            lam, mj, sj = JUMPS if model else (0.0, 0.0, 0.0)
            for t in EXPIRIES:
                for k in STRIKES:
                    rows.append((100.0, k, t, 0.03, 0.0, 0.0, 0, 0, v0, kappa, theta, sigma, rho, lam, mj, sj,
                                 model, 0))
    return np.array(rows, dtype=ope.CF_DTYPE)


def analytic_cf(ope, wm, steps=5, with_cpu=True):
    """GPU analytic prices per second, roofline against FP64, the oracle's
    restatement of the reference pricer on the host cores beside it."""
    import numpy as np

    opts = cf_options(ope)
    n = len(opts)
    ope.cf_price_options(opts)  # warm-up
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        prices, panels, k = ope.cf_price_options(opts, with_panels=True, with_kernel_ms=True)
        kms.append(k)
    wall = (time.perf_counter() - t0) / steps
    kernel_ms = sum(kms) / len(kms)
    flops = sum(wm.cf_flops("bates" if m else "heston", int(p[0]), int(p[1])) for m, p in zip(opts["model"], panels))
    achieved = flops / (kernel_ms * 1e-3) / 1e12
    desc = ("32 Heston parameter sets (numpy seed 11: v0, theta ~ U[0.02, 0.09], kappa ~ U[0.5, 4], sigma ~ "
            "U[0.2, 1], rho ~ U[-0.9, -0.1]) x 41 call strikes 60..140 x 8 expiries 0.1..2, each as Heston and as "
            f"Bates (lambda=0.5, mu_j=-0.1, sigma_j=0.15); S0=100, r=0.03, q=0; {n} options per call")
    out = {"workload": desc, "options": n, "options_per_s": n / wall, "ms_per_call": wall * 1e3,
           "kernel_ms": kernel_ms, "kernel_options_per_s": n / (kernel_ms * 1e-3),
           "mean_panels_per_integral": float(panels.mean()),
           "roofline": {"bound": "valu", "achieved": achieved, "peak": wm.PEAK_TFLOPS["f64"], "unit": "TFLOP/s",
                        "frac": achieved / wm.PEAK_TFLOPS["f64"], "traffic": load_traffic("analytic_cf")},
           "dtype": "f64"}
    print(f"# analytic_cf: {n} options {wall * 1e3:.2f} ms/call, kernel {kernel_ms:.2f} ms",
          file=sys.stderr, flush=True)
    if with_cpu:
        import oracle

        cores = max(1, min(16, os.cpu_count() or 1))
        args = np.stack([opts[f] for f in ("spot", "strike", "expiry", "rate", "div_yield", "v0", "kappa", "theta",
                                           "sigma", "rho", "lambda_", "mu_j", "sigma_j")]
                        + [opts["model"].astype(float), opts["opt_type"].astype(float)], axis=1)
        t0 = time.perf_counter()
        ref = oracle.cf_price_many(args, n_threads=cores)
        cpu_wall = time.perf_counter() - t0
        out["max_abs_diff_vs_oracle"] = float(np.max(np.abs(ref - prices)))
        out["cpu_baseline"] = {"value": n / cpu_wall, "unit": "options/s", "cores": cores, "kind": "port",
                               "sample": f"the same {n} options, C restatement of heston.rs / bates.rs, "
                                         f"{cores} threads, {cpu_wall:.2f} s wall", **host_info()}
    return out


def bsm_iv(ope, wm, steps=5, with_cpu=True, n_bsm=1 << 22, n_iv=1 << 20):
    """BSM price-and-Greeks chain (HBM-bound: a 56-byte contract in, a 64-byte
    result out per option) and an implied-volatility chain on the GPU, with
    the oracle's restatements of bsm.rs / iv.rs on the host cores beside them."""
    import numpy as np

    rng = np.random.default_rng(21)
    c = np.zeros(n_bsm, dtype=ope.CONTRACT_DTYPE)
    c["spot"] = 100.0
    c["strike"] = rng.uniform(50.0, 200.0, n_bsm)
    c["expiry"] = rng.uniform(0.05, 3.0, n_bsm)
    c["rate"] = 0.03
    c["div_yield"] = rng.uniform(0.0, 0.03, n_bsm)
    c["vol"] = rng.uniform(0.1, 0.8, n_bsm)
    c["opt_type"] = rng.integers(0, 2, n_bsm)
    g, _ = ope.bsm_chain(c)  # warm-up
    kms, t0 = [], time.perf_counter()
    for _ in range(steps):
        g, k = ope.bsm_chain(c)
        kms.append(k)
    wall = (time.perf_counter() - t0) / steps
    k_bsm = sum(kms) / len(kms)
    nbytes = n_bsm * (ope.CONTRACT_DTYPE.itemsize + ope.GREEKS_DTYPE.itemsize)
    gbs = nbytes / (k_bsm * 1e-3) / 1e9
    ci, prices = c[:n_iv], g["price"][:n_iv]
    ope.implied_vol_chain(ci, prices)
    kms, t0 = [], time.perf_counter()
    for _ in range(steps):
        vol, found, k = ope.implied_vol_chain(ci, prices)
        kms.append(k)
    wall_iv = (time.perf_counter() - t0) / steps
    k_iv = sum(kms) / len(kms)
    out = {"workload": (f"{n_bsm} BSM contracts (S=100, K ~ U[50, 200], T ~ U[0.05, 3], r=0.03, q ~ U[0, 0.03], "
                        f"vol ~ U[0.1, 0.8], calls and puts, numpy seed 21): price and 7 Greeks; implied vol of "
                        f"the first {n_iv} from their prices"),
           "bsm_options_per_s": n_bsm / (k_bsm * 1e-3), "bsm_kernel_ms": k_bsm, "bsm_ms_per_call": wall * 1e3,
           "bsm_roofline": {"bound": "hbm", "achieved": gbs, "peak": wm.PEAK_HBM_GBS, "unit": "GB/s",
                            "frac": gbs / wm.PEAK_HBM_GBS,
                            "traffic": load_traffic("bsm_chain_greeks") if n_bsm == 1 << 22 else None},
           "iv_options_per_s": n_iv / (k_iv * 1e-3), "iv_kernel_ms": k_iv, "iv_ms_per_call": wall_iv * 1e3,
           "iv_found": int(found.sum()), "dtype": "f64"}
    print(f"# bsm: {n_bsm} options kernel {k_bsm:.3f} ms ({gbs:.0f} GB/s); iv: {n_iv} kernel {k_iv:.3f} ms",
          file=sys.stderr, flush=True)
    if with_cpu:
        import oracle

        cores = max(1, min(16, os.cpu_count() or 1))
        m = 1 << 18
        a = np.column_stack([c[f][:m] for f in ("spot", "strike", "expiry", "rate", "div_yield", "vol")]
                            + [c["opt_type"][:m].astype(float)])
        t0 = time.perf_counter()
        ref = oracle.bsm_greeks_many(a, n_threads=cores)
        t_bsm = time.perf_counter() - t0
        b = np.column_stack([a[:, :5], a[:, 6], ref[:, 0]])
        t0 = time.perf_counter()
        oracle.implied_vol_many(b, n_threads=cores)
        t_iv = time.perf_counter() - t0
        out["max_rel_diff_vs_oracle"] = float(np.max(np.abs(g["price"][:m] - ref[:, 0]) / np.maximum(1.0, ref[:, 0])))
        out["cpu_baseline"] = {"bsm_options_per_s": m / t_bsm, "iv_options_per_s": m / t_iv, "cores": cores,
                               "kind": "port", "sample": f"the first {m} contracts, C restatement of bsm.rs / iv.rs, "
                                                         f"{cores} threads", **host_info()}
    return out


def calibration_eval(ope, steps=5, with_cpu=True):
    """One Levenberg-Marquardt iteration's evaluations (residuals, Jacobian,
    J'J, J'r) for a 41-strike x 8-expiry quote surface, Heston and Bates, on
    the GPU, with the oracle's restatement of calibration.rs beside it."""
    import numpy as np

    gen = ope.HestonParams(0.04, 1.5, 0.04, 0.5, -0.7)
    gen_b = ope.BatesParams(gen, *JUMPS)
    out = {}
    for name, g, p in (("heston", gen, ope.HestonParams(0.05, 2.0, 0.045, 0.4, -0.6)),
                       ("bates", gen_b, ope.BatesParams(ope.HestonParams(0.05, 2.0, 0.045, 0.4, -0.6), 0.4, -0.08,
                                                        0.12))):
        chain = np.zeros(len(EXPIRIES) * len(STRIKES), dtype=ope.CONTRACT_DTYPE)
        chain["spot"] = 100.0
        chain["strike"] = np.tile(STRIKES, len(EXPIRIES))
        chain["expiry"] = np.repeat(EXPIRIES, len(STRIKES))
        chain["rate"] = 0.03
        opts = np.zeros(len(chain), dtype=ope.CF_DTYPE)
        for f in ("spot", "strike", "expiry", "rate", "div_yield", "opt_type"):
            opts[f] = chain[f]
        hp = g.heston if name == "bates" else g
        for f in ("v0", "kappa", "theta", "sigma", "rho"):
            opts[f] = getattr(hp, f)
        if name == "bates":
            opts["lambda_"], opts["mu_j"], opts["sigma_j"], opts["model"] = JUMPS[0], JUMPS[1], JUMPS[2], 1
        px = ope.cf_price_options(opts)
        iv, found, _ = ope.implied_vol_chain(chain, px)
        q = np.zeros(len(chain), dtype=ope.CALIB_DTYPE)
        q["contract"], q["iv_market"], q["weight"] = chain, np.where(found == 1, iv, 0.2), 1.0
        ope.calib_jacobian(q, p)
        t0 = time.perf_counter()
        for _ in range(steps):
            res, jac, jtj, grad = ope.calib_jacobian(q, p)
        wall = (time.perf_counter() - t0) / steps
        dim = jac.shape[1]
        entry = {"quotes": len(q), "parameters": dim, "prices_per_eval": len(q) * (1 + 2 * dim),
                 "ms_per_eval": wall * 1e3}
        if with_cpu:
            import oracle

            rows = np.column_stack([q["contract"][f] for f in ("spot", "strike", "expiry", "rate", "div_yield")]
                                   + [q["contract"]["opt_type"], q["iv_market"], q["weight"]])
            pv = [p.heston.v0, p.heston.kappa, p.heston.theta, p.heston.sigma, p.heston.rho, p.lambda_, p.mu_j,
                  p.sigma_j] if name == "bates" else [p.v0, p.kappa, p.theta, p.sigma, p.rho]
            t0 = time.perf_counter()
            ref = oracle.calib_eval(rows, 1 if name == "bates" else 0, pv)
            cpu = time.perf_counter() - t0
            rel = np.abs(jac - ref[1]) / np.maximum(1.0, np.abs(ref[1]))
            entry["max_rel_diff_vs_oracle"] = float(rel.max())
            # where the BSM vega at the model's implied vol is small, that vol
            # is fixed only to the solver's price tolerance 1e-10 (iv.rs:12):
            # two correct solves may differ by 2e-10 / vega, a differenced
            # column (bump h) by 4e-10 / (vega 2 h)
            vega = np.array([oracle.bsm_greeks(r[0], r[1], r[2], r[3], r[4], rr + r[6], int(r[5]))["vega"]
                             for r, rr in zip(rows, ref[0])])
            ok = vega >= 1e-2
            bumps = np.array([1e-4, 1e-3, 1e-4, 1e-3, 1e-4, 1e-4, 1e-4, 1e-4][:dim])
            bound = 1e-8 * np.maximum(1.0, np.abs(ref[1])) + 4e-10 / (vega[:, None] * 2 * bumps[None, :])
            entry["quotes_vega_ge_1e-2"] = int(ok.sum())
            entry["max_rel_diff_vs_oracle_vega_ge_1e-2"] = float(rel[ok].max())
            entry["all_within_iv_solver_tolerance"] = bool(np.all(np.abs(jac - ref[1]) <= bound))
            entry["cpu_baseline"] = {"ms_per_eval": cpu * 1e3, "cores": 1, "kind": "port",
                                     "sample": "the same evaluation, C restatement of calibration.rs, one thread "
                                               "(the reference's evaluation loop is sequential)", **host_info()}
        out[name] = entry
        print(f"# calibration {name}: {wall * 1e3:.2f} ms per evaluation ({entry['prices_per_eval']} prices)",
              file=sys.stderr, flush=True)
    out["workload"] = ("residuals, Jacobian (central differences; Bates jump columns dP/dparam / vega as the reference's AD), J'J and J'r of 328 quotes (41 strikes 60..140 x 8 expiries "
                       "0.1..2, implied vols of the Heston / Bates prices at v0=theta=0.04, kappa=1.5, sigma=0.5, "
                       "rho=-0.7 [, lambda=0.5, mu_j=-0.1, sigma_j=0.15]) at a parameter vector away from it")
    return out


def serving(ope, seconds):
    """The serving configuration (Poisson arrivals at 2000 req/s) in both
    precisions, plus the same requests all at once (saturated throughput)."""
    srv = {}
    for p, name in ((1, "fp32"), (0, "fp64")):
        for rate, tag in ((2000.0, ""), (math.inf, "_burst")):
            r = serve(ope, p, rate, seconds, 1e-3)
            srv[name + tag] = r
            print(f"# serving {name}{tag}: p50 {r['p50_ms']:.2f} ms p99 {r['p99_ms']:.2f} ms "
                  f"{r['path_steps_per_s']:.3g} path-steps/s", file=sys.stderr, flush=True)
    # a looser target that most requests reach before their cap, so that the
    # adaptive rounds (stop at the target, or post the tiles the SE asks for)
    # decide most results
    r = serve(ope, 1, 2000.0, seconds, 1e-2)
    srv["fp32_target_1e-2"] = r
    print(f"# serving fp32 target 1e-2: p50 {r['p50_ms']:.2f} ms p99 {r['p99_ms']:.2f} ms, met {r['met_target']} "
          f"of {r['requests']}, mean rounds {r['mean_rounds']:.2f}", file=sys.stderr, flush=True)
    srv["workload"] = ("Poisson arrivals at 2000 req/s, requests drawn as the uneven-batch configuration, "
                       "each run until std_error <= 1e-3 x price (cap: its drawn path count); _burst: the "
                       "same number of requests submitted at once; fp32_target_1e-2: the arrivals of fp32 "
                       "with the target 1e-2 x price")
    return srv


def time_mode(run, precision, steps, warmup):
    for _ in range(warmup):
        run(precision)
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        prices, ses, k = run(precision)
        kms.append(k)
    wall = time.perf_counter() - t0
    return prices, ses, wall / steps * 1e3, sum(kms) / len(kms)


def barrier_fp64_cost(ope, steps=10):
    """The QE up-and-out call of the headline workload (1M paths x 365 steps,
    B=130) in FP32-path mode with and without OPE_FLAG_BARRIER_FP64: kernel
    time and price of each (the flag decides near-B paths by their FP64
    path)."""
    hp = ope.HestonParams(**HP)
    cfg = ope.McConfig(n_paths=1_000_000, n_steps=365, seed=42, scheme=ope.VarianceScheme.QuadraticExponential)
    pay = ope.Payoff.UpAndOut(100.0, 130.0, 0.0)
    out = {}
    for flag in (False, True):
        kms, r = [], None
        for i in range(steps + 2):
            r = ope.mc_price_ex(100.0, 1.0, 0.03, 0.0, hp, pay, cfg, None, ope.McOptions(1, barrier_fp64=flag))
            if i >= 2:
                kms.append(r.kernel_ms)
        out["on" if flag else "off"] = {"kernel_ms": sum(kms) / len(kms), "price": r.result.price,
                                        "std_error": r.result.std_error}
    print(f"# barrier_fp64: off {out['off']['kernel_ms']:.3f} ms, on {out['on']['kernel_ms']:.3f} ms",
          file=sys.stderr, flush=True)
    return out


def host_info():
    """The host the CPU baseline ran on: CPU model and logical CPU count."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    import oracle

    return {"cpu_model": model, "nproc": os.cpu_count(),
            "build": oracle.lib().ora_build_flags().decode()}


def cpu_uneven_sample(ope, cores, budget=6e8):
    """The uneven batch's requests in order, each priced by the C restatement
    of mc.rs on `cores` threads, until `budget` path-steps have been done."""
    import oracle

    reqs = draw_uneven_requests(ope, 256, 4)
    done = n_req = 0
    t0 = time.perf_counter()
    for q in reqs:
        hp = oracle.heston(q.params.v0, q.params.kappa, q.params.theta, q.params.sigma, q.params.rho)
        pay = oracle.payoff(q.payoff.kind, q.payoff.strike, int(q.payoff.opt_type), q.payoff.barrier,
                            q.payoff.rebate)
        jp = oracle.jumps(q.lambda_, q.mu_j, q.sigma_j) if q.lambda_ > 0 else None
        oracle.mc_price(q.spot, q.expiry, q.rate, q.div_yield, hp, pay, q.cfg.n_paths, q.cfg.n_steps, q.cfg.seed,
                        True, oracle.EULER, jp, n_threads=cores)
        done += q.cfg.n_paths * q.cfg.n_steps
        n_req += 1
        if done >= budget:
            break
    wall = time.perf_counter() - t0
    return {"value": done / wall, "unit": "path-steps/s", "cores": cores, "kind": "port",
            "sample": f"the first {n_req} of the 256 requests ({done:.3g} path-steps), one mc_price call each; "
                      f"{cores} threads, chunk seeding as mc.rs:159-172; {wall:.2f} s wall", **host_info()}


def cpu_baseline(workload, ope=None):
    """C restatement of mc.rs (oracle/), FP64, on this host's cores, bounded sample."""
    import oracle

    cores = max(1, min(16, os.cpu_count() or 1))
    if workload == "uneven_batch":
        return cpu_uneven_sample(ope, cores)
    hp = oracle.heston(**HP)
    if workload == "bates_surface":
        n_paths, n_steps, scheme, pay, jp = 1_000_000, 252, oracle.EULER, oracle.payoff(), oracle.jumps(*JUMPS)
        sample = (f"Bates Euler European K=100 T=1, {n_paths} paths x {n_steps} steps (a quarter of one of the "
                  "328 per-strike calls the reference makes for the surface)")
    elif workload == "heston_euler":
        n_paths, n_steps, scheme, pay, jp = 100_000, 252, oracle.EULER, oracle.payoff(), None
        sample = f"Heston Euler European, {n_paths} paths x {n_steps} steps (the full workload)"
    else:
        n_paths, n_steps, scheme, pay, jp = 1_000_000, 365, oracle.QE, oracle.payoff(oracle.ASIAN, 100.0), None
        sample = (f"Heston QE arithmetic Asian K=100, {n_paths} paths x {n_steps} steps (the workload's Asian "
                  "simulation in full; the up-and-out one costs the same)")
    t0 = time.perf_counter()
    r = oracle.mc_price(100.0, 1.0, 0.03, 0.0, hp, pay, n_paths, n_steps, 42, True, scheme, jp, n_threads=cores)
    wall = time.perf_counter() - t0
    return {"value": n_paths * n_steps / wall, "unit": "path-steps/s", "cores": cores, "kind": "port",
            "sample": sample + f"; {cores} threads, chunk seeding as mc.rs:159-172; {wall:.2f} s wall, "
                               f"~{wall * cores:.0f} thread-seconds",
            "price": r[0], "std_error": r[1], **host_info()}


def load_traffic(kernel_tag):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(kernel_tag)
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    # untimed calls before the timed ones: the GPU clock ramps over a
    # workload's first ~20-30 calls (QE pair kernels 1.006 ms after 3 calls,
    # 0.974-0.985 after 10, 0.962-0.966 after 30)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--workload", default="qe_asian_uao",
                    choices=["qe_asian_uao", "bates_surface", "heston_euler", "uneven_batch", "serving",
                             "analytic_cf"])
    ap.add_argument("--serve-seconds", type=float, default=2.0, help="arrival window of the serving measurement")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving measurement")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the secondary Bates-surface and uneven-batch measurements")
    ap.add_argument("--profile", action="store_true", help="time FP32-path mode only (for rocprofv3 runs)")
    ap.add_argument("--precision", type=int, default=1, choices=[0, 1],
                    help="with --profile: 1 = FP32-path (default), 0 = FP64")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a plain 1 GiB read + 1 GiB write elementwise kernel (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this engine's path has no multi-GPU mode in this build "
              "(the design shards (request, pair-range) tiles with one 16-byte-per-strike reduction)",
              file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        import torch

        n = 1 << 28  # 2^28 float32 = 1 GiB read and 1 GiB written per launch
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            torch.add(a, 1.0, out=b)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration": {"bytes_read_per_launch": 4 * n, "bytes_written_per_launch": 4 * n,
                                               "launches": 3}}))
        return

    import opemc as ope
    from opemc import work_model as wm

    info = ope.device_info()
    print(f"# device {info['name']} {info['arch']} CUs={info['compute_units']} clock={info['clock_mhz']} MHz",
          file=sys.stderr, flush=True)
    if args.workload == "serving":
        print(json.dumps({"serving": serving(ope, args.serve_seconds)}))
        return
    if args.workload == "analytic_cf":
        print(json.dumps({"analytic_cf": analytic_cf(ope, wm, steps=args.steps, with_cpu=not args.no_cpu),
                          "bsm_iv": bsm_iv(ope, wm, steps=args.steps, with_cpu=not args.no_cpu),
                          "calibration_eval": calibration_eval(ope, steps=args.steps, with_cpu=not args.no_cpu)}))
        return
    desc, conf, run, path_steps, flops = make_workload(ope, wm, args.workload)

    if args.profile:
        p, se, ms, k = time_mode(run, args.precision, args.steps, args.warmup)
        print(json.dumps({"workload": args.workload, "precision": args.precision, "ms_per_step": ms,
                          "kernel_ms": k, "path_steps_per_step": path_steps}))
        return
    p32, se32, ms32, k32 = time_mode(run, 1, args.steps, args.warmup)
    print(f"# fp32 path: {ms32:.3f} ms/step wall, {k32:.3f} ms kernel", file=sys.stderr, flush=True)
    p64, se64, ms64, k64 = time_mode(run, 0, max(3, args.steps // 2), 1)
    print(f"# fp64: {ms64:.3f} ms/step wall, {k64:.3f} ms kernel", file=sys.stderr, flush=True)

    err_se = [abs(a - b) / s for a, b, s in zip(p32, p64, se64) if s > 0]
    achieved32 = flops / (k32 * 1e-3) / 1e12
    achieved64 = flops / (k64 * 1e-3) / 1e12
    line = {
        "metric": METRIC,
        "value": path_steps / (ms32 * 1e-3),
        "unit": "path-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (Philox4x32-10 draws keyed by seed and pair)",
        "config": dict(workload=desc, precision="fp32-path, fp64 log-price sum and reduction", **conf),
        "roofline": {"bound": "valu", "achieved": achieved32, "peak": wm.PEAK_TFLOPS["f32"], "unit": "TFLOP/s",
                     "frac": achieved32 / wm.PEAK_TFLOPS["f32"],
                     "traffic": load_traffic(f"{args.workload}_fp32")},
        "kernel_ms": k32,
        "path_steps_per_step": path_steps,
        "fp64": {"value": path_steps / (ms64 * 1e-3), "ms_per_step": ms64, "kernel_ms": k64,
                 "roofline": {"bound": "valu", "achieved": achieved64, "peak": wm.PEAK_TFLOPS["f64"],
                              "unit": "TFLOP/s", "frac": achieved64 / wm.PEAK_TFLOPS["f64"]}},
        "fp32_over_fp64": ms64 / ms32,
        "fp32_price_err_max_se": max(err_se) if err_se else None,
        "prices_fp32": p32[:4],
        "prices_fp64": p64[:4],
        "std_errors": se64[:4],
        "serving": None,  # filled below
    }
    if args.workload == "qe_asian_uao" and not args.no_extra:
        for extra in ("bates_surface", "uneven_batch", "heston_euler"):
            line[extra] = secondary(ope, wm, extra)
            if not args.no_cpu:
                cb = cpu_baseline(extra, ope)
                # the whole workload on the host at the sampled rate; the
                # reference prices the surface as 328 separate (expiry, strike)
                # calls, each simulating its own 4M paths: 41x the path-steps
                # of the shared-path surface
                resim = len(STRIKES) if extra == "bates_surface" else 1
                cb["extrapolated_workload_s"] = line[extra]["path_steps"] * resim / cb["value"]
                line[extra]["cpu_baseline"] = cb
        line["barrier_fp64_flag"] = barrier_fp64_cost(ope)
        line["analytic_cf"] = analytic_cf(ope, wm, with_cpu=not args.no_cpu)
        line["bsm_iv"] = bsm_iv(ope, wm, with_cpu=not args.no_cpu)
        line["calibration_eval"] = calibration_eval(ope, with_cpu=not args.no_cpu)
    if not args.no_serve:
        line["serving"] = serving(ope, args.serve_seconds)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args.workload, ope)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line))


if __name__ == "__main__":
    main()
