"""Writes the golden fixtures of this directory from the oracle (the C
restatement of the reference, oracle/): small inputs and the outputs the
reference's algorithm gives for them.  The reference ships no golden vectors
for this path (it is Rust, with no data files), so these pin the oracle
against accidental change and give the GPU tests fixed expectations.

    python tests/golden/make_golden.py

Fixtures (numpy .npz, no pickled objects):
  replay.npz  run_path (mc.rs:210-278) on the chunk-seeded SmallRng draws of
              mc.rs:159-180 (seed 42, 32 pairs x 12 steps, one thread), at the
              MC workloads' parameters, for Euler / QE x Heston / Bates x
              European / Asian / up-and-out: draws in, per-path payoffs and
              states (ln S_T, v_T, sum S, max S) out;
  cf.npz      heston_price / bates_price (heston.rs:61-72, bates.rs:29-39) on
              the 41-strike x 8-expiry call grid at the same parameters;
  philox.npz  the engine's Philox-mode draws (normals, jump sums, Poisson
              counts) for 16 pairs around the counter's 2^32 boundary, 19
              steps, lambda 0 / 0.5 / 30.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as ora  # noqa: E402

S0, K, T, R, Q = 100.0, 100.0, 1.0, 0.03, 0.0
HP = dict(v0=0.04, kappa=1.5, theta=0.04, sigma=0.5, rho=-0.7)
JUMP = (0.5, -0.1, 0.15)
BARRIER = 130.0
STRIKES = np.arange(60.0, 140.0 + 1e-9, 2.0)
EXPIRIES = np.array([0.1, 0.25, 0.5, 0.75, 1.0, 1.25, 1.5, 2.0])
REPLAY_PAIRS, REPLAY_STEPS, SEED = 32, 12, 42
PHILOX_PAIRS = np.concatenate([np.arange(8), 2**32 - 4 + np.arange(8)]).astype(np.uint64)


def replay_cases():
    for scheme in (ora.EULER, ora.QE):
        for lam in (None, JUMP[0]):
            for kind in (ora.EUROPEAN, ora.ASIAN, ora.UP_AND_OUT):
                yield scheme, lam, kind


def replay_case(scheme, lam, kind):
    jp = ora.jumps(lam, *JUMP[1:]) if lam is not None else ora.jumps()
    za, zb, js = ora.gen_draws(SEED, REPLAY_PAIRS, REPLAY_STEPS, T, jp, n_threads=1)
    pay = ora.payoff(kind, K, ora.CALL, BARRIER if kind == ora.UP_AND_OUT else 0.0, 0.0)
    payoff, state = ora.run_paths(S0, T, R, Q, ora.heston(**HP), pay, REPLAY_STEPS, scheme, True, za, zb, js, jp)
    return za, zb, js, payoff, state


def key(scheme, lam, kind):
    return f"s{scheme}_l{0 if lam is None else 1}_k{kind}"


def cf_grid():
    hp = ora.heston(**HP)
    heston = np.array([[ora.heston_price(S0, k, t, R, Q, hp) for k in STRIKES] for t in EXPIRIES])
    bates = np.array([[ora.bates_price(S0, k, t, R, Q, hp, *JUMP) for k in STRIKES] for t in EXPIRIES])
    return heston, bates


def philox_case(lam):
    return ora.philox_draws(SEED, PHILOX_PAIRS, 19, dt=1.0 / 19, lam=lam, mu_j=JUMP[1], sigma_j=JUMP[2])


def main():
    out = {}
    for c in replay_cases():
        za, zb, js, payoff, state = replay_case(*c)
        k = key(*c)
        out[k + "_za"], out[k + "_zb"], out[k + "_js"] = za, zb, js
        out[k + "_payoff"], out[k + "_state"] = payoff, state
    np.savez_compressed(os.path.join(HERE, "replay.npz"), **out)
    heston, bates = cf_grid()
    np.savez_compressed(os.path.join(HERE, "cf.npz"), strikes=STRIKES, expiries=EXPIRIES, heston=heston,
                        bates=bates)
    ph = {}
    for i, lam in enumerate((0.0, 0.5, 30.0)):
        za, zb, js, nn = philox_case(lam)
        ph[f"l{i}_za"], ph[f"l{i}_zb"], ph[f"l{i}_js"], ph[f"l{i}_n"] = za, zb, js, nn
    np.savez_compressed(os.path.join(HERE, "philox.npz"), pairs=PHILOX_PAIRS, **ph)
    print("wrote", ", ".join(sorted(f for f in os.listdir(HERE) if f.endswith(".npz"))))


if __name__ == "__main__":
    main()
