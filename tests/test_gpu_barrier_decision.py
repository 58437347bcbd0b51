"""FP32-path up-and-out with OPE_FLAG_BARRIER_FP64: the knock-out decision
near the barrier is FP64's.

A path whose FP32 running maximum lands within a margin (2^-16 in log2
units) of log2(B/S0) is simulated again by the FP64 path on the same draws
and takes that path's state, knock-out flag and terminal value.  Checked in
replay mode, where per-path states come back: the barrier is set to the
FP64 running maximum of a chosen path, so the FP64 path knocks out exactly
at B (running_max >= B, mc.rs:91, 218, 272) while the FP32 maximum of the
same path differs from it by rounding; the FP32 run must decide, and
price, that path as FP64 does, bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HP = dict(v0=0.04, kappa=1.5, theta=0.04, sigma=0.5, rho=-0.7)


def _run(ope, precision, barrier, cfg, draws, jumps, flag=True):
    pay = ope.Payoff.UpAndOut(100.0, barrier, rebate=5.0)
    r = ope.mc_price_ex(100.0, 1.0, 0.03, 0.0, ope.HestonParams(**HP), pay, cfg, jumps,
                        ope.McOptions(precision, replay=draws, want_paths=True, barrier_fp64=flag))
    return r.path_payoffs, r.path_states


@pytest.mark.parametrize("scheme", [0, 1])
@pytest.mark.parametrize("lam", [0.0, 0.5])
def test_fp32_decides_exact_barrier_hits_as_fp64(ope, scheme, lam):
    rng = np.random.default_rng(11 + scheme)
    n_pairs, n_steps = 256, 120
    za, zb = rng.standard_normal((n_steps, n_pairs)), rng.standard_normal((n_steps, n_pairs))
    js = np.where(rng.random((n_steps, n_pairs)) < 0.02, rng.normal(-0.1, 0.15, (n_steps, n_pairs)), 0.0)
    jumps = (lam, -0.1, 0.15) if lam > 0 else None
    draws = (za, zb, js if lam > 0 else None)
    cfg = ope.McConfig(n_paths=2 * n_pairs, n_steps=n_steps, seed=1, scheme=scheme)
    _, free = _run(ope, 0, 1e300, cfg, draws, jumps)          # FP64 running maxima, no barrier
    maxima = free[..., 3]
    chosen = [(p, s) for p, s in zip(rng.choice(n_pairs, 12, replace=False), [0, 1] * 6) if maxima[p, s] > 100.0]
    print(f"scheme {scheme} lam {lam}: {len(chosen)} barriers at a path's FP64 maximum", flush=True)
    assert len(chosen) >= 6
    for p, s in chosen:
        b = float(maxima[p, s])
        pay64, st64 = _run(ope, 0, b, cfg, draws, jumps)
        pay32, st32 = _run(ope, 1, b, cfg, draws, jumps)
        assert st64[p, s, 3] == b                        # FP64: the path touches B and knocks out
        assert pay64[p, s] == pytest.approx(5.0 * np.exp(-0.03))
        assert pay32[p, s] == pay64[p, s]                # FP32 decides it as FP64 does ...
        assert np.array_equal(st32[p, s], st64[p, s], equal_nan=True)    # ... from the FP64 path's state
    # elsewhere the two runs agree on the knock-outs up to paths far from B
    # whose trajectories diverged (none at these sizes)
    ko32 = np.isclose(pay32, 5.0 * np.exp(-0.03), rtol=0, atol=1e-12)
    ko64 = np.isclose(pay64, 5.0 * np.exp(-0.03), rtol=0, atol=1e-12)
    assert (ko32 != ko64).sum() == 0


def test_without_the_flag_the_fp32_path_decides(ope):
    # the default: every sibling's state comes from the FP32 path (the
    # running maximum of a chosen path is then FP32's, not FP64's)
    rng = np.random.default_rng(19)
    n_pairs, n_steps = 64, 60
    za, zb = rng.standard_normal((n_steps, n_pairs)), rng.standard_normal((n_steps, n_pairs))
    cfg = ope.McConfig(n_paths=2 * n_pairs, n_steps=n_steps, seed=1, scheme=1)
    _, free = _run(ope, 0, 1e300, cfg, (za, zb, None), None)
    p = int(np.argmax(free[:, 0, 3]))
    b = float(free[p, 0, 3])
    _, on = _run(ope, 1, b, cfg, (za, zb, None), None, flag=True)
    _, off = _run(ope, 1, b, cfg, (za, zb, None), None, flag=False)
    assert on[p, 0, 3] == b
    assert off[p, 0, 3] != b and abs(off[p, 0, 3] / b - 1) < 1e-5
