"""The uneven-batch and serving workloads at their full size and load.

* 256 independent Heston/Bates requests (v0, theta ~ U[0.01, 0.16], kappa ~
  U[0.5, 5], xi ~ U[0.1, 1], rho ~ U[-0.95, 0], lambda ~ U[0, 1], payoff
  uniform over European / Asian / up-and-out B=130, paths log-uniform in
  [1e4, 2e6], steps uniform in [50, 500], Euler, T = 1) in one ope_mc_batch
  per precision: every FP32-path price within 0.1 standard errors of the FP64
  price on the same draws, and the eight smallest requests (all three payoffs
  among them) within 3 combined standard errors of the CPU oracle's
  mc_bates at the same path count (mc.rs:107-113, 148-207).
* The server under Poisson arrivals at 2000 requests/s drawn the same way,
  each run to SE <= 1e-3 x price (capped at its drawn path count), with a
  64 GiB admission budget, in both precisions: every served result equals
  ope_mc_bates_ex at the path count the server ran, bit for bit, and p50/p99
  latency is reported.
"""
import math

import numpy as np
import pytest

import bench

pytestmark = pytest.mark.gpu

WAIT_MS = 120_000.0


@pytest.fixture(scope="module")
def uneven(ope):
    reqs = bench.draw_uneven_requests(ope, 256, 4)
    out = {}
    for precision in (0, 1):
        print(f"256-request batch, precision={precision}", flush=True)
        res, _, _ = ope.mc_batch(reqs, precision)
        out[precision] = res
    return reqs, out


def test_uneven_batch_fp32_within_a_tenth_of_a_standard_error(uneven):
    reqs, out = uneven
    assert len(reqs) == 256
    errs = [abs(a.price - b.price) / b.std_error for a, b in zip(out[1], out[0])]
    print(f"max |fp32 - fp64| = {max(errs):.4f} SE", flush=True)
    assert max(errs) <= 0.1
    for a, b in zip(out[1], out[0]):
        assert abs(a.std_error / b.std_error - 1.0) < 0.01


def test_uneven_batch_smallest_requests_match_the_oracle(ora, uneven):
    reqs, out = uneven
    order = sorted(range(len(reqs)), key=lambda i: reqs[i].cfg.n_paths * reqs[i].cfg.n_steps)
    chosen = []
    for kind in (0, 1, 2):          # the two smallest of each payoff, then the smallest of the rest
        chosen += [i for i in order if reqs[i].payoff.kind == kind][:2]
    chosen += [i for i in order if i not in chosen][:8 - len(chosen)]
    assert len(chosen) == 8 and {reqs[i].payoff.kind for i in chosen} == {0, 1, 2}
    for i in chosen:
        q = reqs[i]
        p = q.params
        pay = ora.payoff(q.payoff.kind, q.payoff.strike, int(q.payoff.opt_type), q.payoff.barrier, q.payoff.rebate)
        c = ora.mc_price(q.spot, q.expiry, q.rate, q.div_yield, ora.heston(p.v0, p.kappa, p.theta, p.sigma, p.rho),
                         pay, q.cfg.n_paths, q.cfg.n_steps, q.cfg.seed, True, 0,
                         ora.jumps(q.lambda_, q.mu_j, q.sigma_j), n_threads=8)
        for precision in (0, 1):
            g = out[precision][i]
            z = abs(g.price - c[0]) / math.hypot(g.std_error, c[1])
            print(f"request {i}: kind {q.payoff.kind} paths {q.cfg.n_paths} steps {q.cfg.n_steps} "
                  f"precision {precision}: {g.price:.5f} vs {c[0]:.5f}, z = {z:.2f}", flush=True)
            assert z < 3.0


@pytest.mark.parametrize("precision", [1, 0])
def test_serving_at_2000_requests_per_second(ope, precision):
    target = 1e-3
    arrivals = bench.poisson_arrivals(2000.0, 0.3, 5)
    reqs = bench.draw_uneven_requests(ope, len(arrivals), 6)
    print(f"serving {len(reqs)} requests at 2000/s, precision={precision}", flush=True)
    with ope.Server(precision, max_requests=4096, memory_budget_bytes=64 << 30, idle_exit_ms=200.0) as srv:
        tickets, t0 = bench.submit_on_schedule(srv, arrivals, reqs, target)
        got = [srv.wait(t, WAIT_MS) for t in tickets]
        st = srv.stats()
    lat = np.array([g.latency_ms for g in got])
    met = sum(1 for g in got if g.stop_reason == ope.SERVER_STOP_TARGET)
    print(f"p50 {np.percentile(lat, 50):.2f} ms, p99 {np.percentile(lat, 99):.2f} ms, "
          f"{met} met the target, peak active {st['peak_active']}", flush=True)
    assert st["completed"] == len(reqs) and st["peak_active"] > 1
    for q, g in zip(reqs, got):
        assert g.stop_reason in (ope.SERVER_STOP_TARGET, ope.SERVER_STOP_ALL_PATHS)
        if g.stop_reason == ope.SERVER_STOP_TARGET:
            assert g.result.std_error <= target * abs(g.result.price)
        else:
            assert g.paths == q.cfg.n_paths
    print("checking every result against the one-shot call", flush=True)
    for q, g in zip(reqs, got):
        cfg = ope.McConfig(n_paths=g.paths, n_steps=q.cfg.n_steps, seed=q.cfg.seed)
        one = ope.mc_price_ex(q.spot, q.expiry, q.rate, q.div_yield, q.params, q.payoff, cfg,
                              (q.lambda_, q.mu_j, q.sigma_j), ope.McOptions(precision))
        assert one.result == g.result
