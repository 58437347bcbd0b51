"""Persistent-kernel serving loop: results equal the batch path bit for bit,
a request's result is the same alone and in a mix of concurrent requests
(also when it stops at a target standard error), admission respects the
memory budget, and the kernel exits when idle and is relaunched on demand."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WAIT_MS = 120_000.0


def uneven_requests(ope, n, seed, max_paths=400_000, min_steps=20, max_steps=120):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        v0, theta = rng.uniform(0.01, 0.16, 2)
        hp = ope.HestonParams(v0, rng.uniform(0.5, 5.0), theta, rng.uniform(0.1, 1.0), rng.uniform(-0.95, 0.0))
        kind = int(rng.integers(0, 3))
        pay = [ope.Payoff.European(100.0), ope.Payoff.AsianArithmetic(100.0), ope.Payoff.UpAndOut(100.0, 130.0)][kind]
        paths = int(np.exp(rng.uniform(np.log(1e4), np.log(max_paths)))) // 2 * 2
        cfg = ope.McConfig(n_paths=paths, n_steps=int(rng.integers(min_steps, max_steps + 1)), seed=500 + i,
                           scheme=int(rng.integers(0, 2)))
        lam = float(rng.uniform(0.0, 1.0)) if rng.random() < 0.5 else 0.0
        out.append(ope.McRequest(100.0, 1.0, 0.03, 0.0, hp, pay, cfg, lam, -0.1, 0.15))
    return out


@pytest.mark.parametrize("precision", [0, 1])
def test_full_runs_equal_the_batch_path(ope, precision):
    reqs = uneven_requests(ope, 16, 3)
    print(f"batch reference precision={precision}", flush=True)
    batch, _, _ = ope.mc_batch(reqs, precision)
    print("serving", flush=True)
    with ope.Server(precision, idle_exit_ms=50.0) as srv:
        tickets = [srv.submit(q) for q in reqs]
        got = [srv.wait(t, WAIT_MS) for t in tickets]
        st = srv.stats()
    print("stats", st, flush=True)
    for g, b, q in zip(got, batch, reqs):
        assert g.result == b
        assert g.paths == q.cfg.n_paths
        assert g.path_steps == q.cfg.n_paths * q.cfg.n_steps
        assert g.rounds == 1
    assert st["completed"] == len(reqs)
    assert st["kernel_launches"] >= 1


@pytest.mark.parametrize("precision", [0, 1])
def test_adaptive_result_is_the_same_alone_and_in_the_mix(ope, precision):
    hp = ope.HestonParams(0.04, 1.5, 0.04, 0.5, -0.7)
    probe = ope.McRequest(100.0, 1.0, 0.03, 0.0, hp, ope.Payoff.European(100.0),
                          ope.McConfig(n_paths=1_800_000, n_steps=60, seed=77), 0.5, -0.1, 0.15)
    target = 2e-3
    print(f"probe alone precision={precision}", flush=True)
    with ope.Server(precision, idle_exit_ms=50.0, initial_round_pairs=16384) as srv:
        alone = srv.wait(srv.submit(probe, target), WAIT_MS)
    print("probe in a mix", alone, flush=True)
    others = uneven_requests(ope, 40, 9)
    with ope.Server(precision, idle_exit_ms=50.0, initial_round_pairs=16384) as srv:
        tickets = [srv.submit(q, 1e-3) for q in others[:20]]
        t_probe = srv.submit(probe, target)
        time.sleep(0.002)
        tickets += [srv.submit(q, 1e-3) for q in others[20:]]
        mixed = srv.wait(t_probe, WAIT_MS)
        for t in tickets:
            srv.wait(t, WAIT_MS)
        st = srv.stats()
    assert mixed.result == alone.result
    assert mixed.paths == alone.paths and mixed.rounds == alone.rounds
    assert alone.rounds >= 2
    assert alone.result.std_error <= target * abs(alone.result.price) or alone.paths == probe.cfg.n_paths
    assert st["peak_active"] > 1
    # ... and equals the one-shot call with the path count the server stopped at
    cfg = ope.McConfig(n_paths=alone.paths, n_steps=60, seed=77)
    one = ope.mc_price_ex(100.0, 1.0, 0.03, 0.0, hp, ope.Payoff.European(100.0), cfg, (0.5, -0.1, 0.15),
                          ope.McOptions(precision))
    assert one.result == alone.result


def test_budget_limits_residency(ope):
    reqs = uneven_requests(ope, 10, 11)
    for q in reqs:
        q.cfg.n_paths = 50_000
    # footprint per request: tiles x (16 B slot + 16 B ring item) + record + completion
    one = ((50_000 // 2 + 255) // 256) * 32 + 1024
    budget = int(2.5 * one)
    with ope.Server(1, memory_budget_bytes=budget, idle_exit_ms=50.0) as srv:
        tickets = [srv.submit(q) for q in reqs]
        for t in tickets:
            srv.wait(t, WAIT_MS)
        st = srv.stats()
    assert st["completed"] == len(reqs)
    assert st["peak_resident_bytes"] <= budget
    assert st["peak_active"] <= 2


def test_idle_exit_and_relaunch(ope):
    q = uneven_requests(ope, 1, 13)[0]
    with ope.Server(1, idle_exit_ms=20.0) as srv:
        a = srv.wait(srv.submit(q), WAIT_MS)
        time.sleep(0.5)   # the kernel drains after 20 ms without work
        b = srv.wait(srv.submit(q), WAIT_MS)
        st = srv.stats()
    assert a.result == b.result
    assert st["kernel_launches"] >= 2


def test_rejections(ope):
    with ope.Server(1, memory_budget_bytes=4096, idle_exit_ms=20.0) as srv:
        big = uneven_requests(ope, 1, 17)[0]
        big.cfg.n_paths = 1_000_000
        with pytest.raises(ope.OpeError) as e:
            srv.submit(big)
        assert e.value.status == ope.OPE_ERR_OVER_BUDGET
        bad = uneven_requests(ope, 1, 17)[0]
        bad.cfg.n_paths = 1
        with pytest.raises(ope.OpeError) as e:
            srv.submit(bad)
        assert e.value.status == ope.OPE_ERR_INVALID_ARGUMENT


def _served_equals_single_call(ope, q, got, precision):
    """A served result equals ope_mc_bates_ex at the path count the server ran."""
    cfg = ope.McConfig(n_paths=got.paths, n_steps=q.cfg.n_steps, seed=q.cfg.seed, antithetic=q.cfg.antithetic,
                       scheme=q.cfg.scheme)
    one = ope.mc_price_ex(q.spot, q.expiry, q.rate, q.div_yield, q.params, q.payoff, cfg,
                          (q.lambda_, q.mu_j, q.sigma_j), ope.McOptions(precision))
    return one.result == got.result


@pytest.mark.parametrize("precision", [0, 1])
def test_one_slot_reused_by_successive_requests(ope, precision):
    # A workgroup keeps the record of its previous item's request and reads
    # the record again only for another request.  Through one slot, twelve
    # different requests follow each other (each adaptive, so a request's
    # later rounds reuse the record its earlier rounds loaded, while every
    # new request in the slot must load its own): each result equals the
    # one-shot call at the path count served, bit for bit.
    reqs = uneven_requests(ope, 12, 41, max_paths=200_000)
    print(f"one slot, {len(reqs)} requests, precision={precision}", flush=True)
    with ope.Server(precision, max_requests=1, idle_exit_ms=50.0, initial_round_pairs=4096) as srv:
        tickets = [srv.submit(q, 1e-9) for q in reqs]
        got = [srv.wait(t, WAIT_MS) for t in tickets]
        st = srv.stats()
    assert st["completed"] == len(reqs) and st["peak_active"] == 1
    assert sum(g.rounds for g in got) > len(reqs)
    for q, g in zip(reqs, got):
        assert g.slot == 0
        assert _served_equals_single_call(ope, q, g, precision), (q, g)


@pytest.mark.parametrize("precision", [0, 1])
def test_adaptive_requests_whose_caps_exceed_the_ring(ope, precision):
    # Six adaptive requests of 30 tiles each (15360 paths) against a 64-tile
    # ring: a request reserves its whole cap at admission, so at most two are
    # resident, and their second rounds (22 tiles each) never overwrite an
    # item a worker has not read.  Target 1e-9 is never met: every request
    # runs its first round of 8 tiles, then the rest of its cap.
    reqs = uneven_requests(ope, 6, 21)
    for q in reqs:
        q.cfg.n_paths = 2 * 30 * 256
    print(f"ring of 64 tiles, precision={precision}", flush=True)
    with ope.Server(precision, idle_exit_ms=50.0, ring_items=64, initial_round_pairs=2048) as srv:
        tickets = [srv.submit(q, 1e-9) for q in reqs]
        got = [srv.wait(t, WAIT_MS) for t in tickets]
        st = srv.stats()
    print("stats", st, flush=True)
    assert st["peak_active"] == 2
    for q, g in zip(reqs, got):
        assert g.paths == q.cfg.n_paths and g.rounds == 2
        assert g.stop_reason == ope.SERVER_STOP_ALL_PATHS
        assert _served_equals_single_call(ope, q, g, precision)


def test_request_larger_than_the_ring_is_rejected(ope):
    with ope.Server(1, idle_exit_ms=20.0, ring_items=64) as srv:
        q = uneven_requests(ope, 1, 23)[0]
        q.cfg.n_paths = 2 * 65 * 256
        with pytest.raises(ope.OpeError) as e:
            srv.submit(q)
        assert e.value.status == ope.OPE_ERR_INVALID_ARGUMENT


def _long_request(ope, seed=91):
    # FP64: 2M paths x 4000 steps, about 35 ms on the device
    hp = ope.HestonParams(0.04, 1.5, 0.04, 0.5, -0.7)
    return ope.McRequest(100.0, 1.0, 0.03, 0.0, hp, ope.Payoff.European(100.0),
                         ope.McConfig(n_paths=2_000_000, n_steps=4000, seed=seed), 0.5, -0.1, 0.15)


def test_cancel_queued_and_running_requests(ope):
    others = uneven_requests(ope, 2, 25)
    print("one slot: a long request runs, two wait", flush=True)
    with ope.Server(0, max_requests=1, idle_exit_ms=50.0) as srv:
        t_long = srv.submit(_long_request(ope))
        t_queued = srv.submit(others[0])
        t_last = srv.submit(others[1])
        srv.cancel(t_queued)   # still queued behind the long request
        srv.cancel(t_long)     # queued or running: fails at its round's end
        for t in (t_queued, t_long):
            with pytest.raises(ope.OpeError) as e:
                srv.wait(t, WAIT_MS)
            assert e.value.status == ope.OPE_ERR_CANCELLED
        last = srv.wait(t_last, WAIT_MS)
        with pytest.raises(ope.OpeError) as e:
            srv.cancel(t_last)   # already finished and collected
        assert e.value.status == ope.OPE_ERR_NOT_FOUND
        st = srv.stats()
    assert _served_equals_single_call(ope, others[1], last, 0)
    assert st["completed"] == 1


def test_deadline_while_queued(ope):
    q = uneven_requests(ope, 1, 27)[0]
    with ope.Server(0, max_requests=1, idle_exit_ms=50.0) as srv:
        t_long = srv.submit(_long_request(ope))
        t_late = srv.submit(q, 0.0, deadline_ms=2.0)   # cannot be admitted before the long one ends
        with pytest.raises(ope.OpeError) as e:
            srv.wait(t_late, WAIT_MS)
        assert e.value.status == ope.OPE_ERR_DEADLINE
        srv.wait(t_long, WAIT_MS)


def test_deadline_stops_an_adaptive_request_at_a_round_boundary(ope):
    # First round: 2^19 pairs x 4000 steps in FP64 (about 17 ms) against a
    # 5 ms deadline and an unreachable target: the request stops after that
    # round and returns the estimate from its 2^20 paths, which equals the
    # one-shot call at that path count.
    q = _long_request(ope, seed=93)
    with ope.Server(0, idle_exit_ms=50.0, initial_round_pairs=1 << 19) as srv:
        g = srv.wait(srv.submit(q, 1e-9, deadline_ms=5.0), WAIT_MS)
    print(g, flush=True)
    assert g.stop_reason == ope.SERVER_STOP_DEADLINE
    assert g.rounds == 1 and g.paths == 1 << 20
    assert _served_equals_single_call(ope, q, g, 0)


def test_watchdog_fails_outstanding_requests(ope):
    # A watchdog of 1 ns fires as soon as tiles are outstanding: the host
    # fails every request with OPE_ERR_TIMEOUT and tells the kernel to drain;
    # the kernel itself runs its published tiles and exits normally.
    reqs = uneven_requests(ope, 3, 29)
    with ope.Server(1, idle_exit_ms=50.0, watchdog_ms=1e-6) as srv:
        tickets = [srv.submit(q) for q in reqs]
        for t in tickets:
            with pytest.raises(ope.OpeError) as e:
                srv.wait(t, WAIT_MS)
            assert e.value.status == ope.OPE_ERR_TIMEOUT
        with pytest.raises(ope.OpeError) as e:
            srv.submit(reqs[0])
        assert e.value.status == ope.OPE_ERR_TIMEOUT
