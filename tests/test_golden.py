"""Golden fixtures (tests/golden/, written by make_golden.py from the oracle):
the oracle and the numpy Philox specification must keep reproducing them, and
the GPU must match them.  The reference ships no golden vectors for this path,
so the fixtures pin the oracle's own outputs: a change to the restatement
that moves a per-path payoff, an analytic price or a Philox-mode draw fails
here first."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, name))  # allow_pickle stays False


@pytest.fixture(scope="module")
def mk():
    import importlib.util

    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLDEN, "make_golden.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_oracle_reproduces_the_replay_fixtures(ora, mk):
    g = load("replay.npz")
    for case in mk.replay_cases():
        k = mk.key(*case)
        za, zb, js, payoff, state = mk.replay_case(*case)
        # the chunk-seeded draw stream is integer arithmetic and Box-Muller in FP64
        np.testing.assert_allclose(za, g[k + "_za"], rtol=1e-15, atol=0, err_msg=k)
        np.testing.assert_allclose(zb, g[k + "_zb"], rtol=1e-15, atol=0, err_msg=k)
        np.testing.assert_allclose(js, g[k + "_js"], rtol=1e-15, atol=1e-300, err_msg=k)
        np.testing.assert_allclose(payoff, g[k + "_payoff"], rtol=1e-13, atol=1e-13, err_msg=k)
        np.testing.assert_allclose(state, g[k + "_state"], rtol=1e-13, atol=1e-300, err_msg=k)


def test_replay_fixtures_exercise_every_branch():
    g = load("replay.npz")
    # QE with the Feller condition violated reaches the exponential branch's point mass
    assert any((g[k][..., 1] == 0.0).any() for k in g.files if k.startswith("s1_") and k.endswith("_state"))
    # some up-and-out paths knock out, some do not
    ko = g["s0_l1_k2_state"][..., 3] >= 130.0
    assert ko.any() and not ko.all()
    # the Bates cases carry jumps
    assert (g["s0_l1_k0_js"] != 0.0).any() and not (g["s0_l0_k0_js"] != 0.0).any()


def test_oracle_reproduces_the_cf_fixture(ora, mk):
    g = load("cf.npz")
    heston, bates = mk.cf_grid()
    np.testing.assert_allclose(heston, g["heston"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(bates, g["bates"], rtol=0, atol=1e-12)
    # the grid's known anchor (heston.rs:296-314 parameters aside): at T = 1,
    # K = 100 the Heston call sits in the usual range for these parameters
    i, j = list(g["expiries"]).index(1.0), list(g["strikes"]).index(100.0)
    assert 7.0 < g["heston"][i, j] < 11.0 and g["bates"][i, j] > g["heston"][i, j]


def test_numpy_philox_spec_reproduces_the_fixture(ora, mk):
    g = load("philox.npz")
    assert np.array_equal(g["pairs"], mk.PHILOX_PAIRS)
    for i, lam in enumerate((0.0, 0.5, 30.0)):
        za, zb, js, nn = mk.philox_case(lam)
        np.testing.assert_allclose(za, g[f"l{i}_za"], rtol=1e-14, atol=1e-15)
        np.testing.assert_allclose(zb, g[f"l{i}_zb"], rtol=1e-14, atol=1e-15)
        np.testing.assert_allclose(js, g[f"l{i}_js"], rtol=1e-14, atol=1e-15)
        assert np.array_equal(nn, g[f"l{i}_n"])
    assert g["l2_n"].max() >= 3 and (g["l1_n"] > 0).any()


@pytest.mark.gpu
def test_gpu_replay_matches_the_fixtures(ope, mk):
    g = load("replay.npz")
    for scheme, lam, kind in mk.replay_cases():
        k = mk.key(scheme, lam, kind)
        print(f"golden replay {k}", flush=True)
        za, zb, js = g[k + "_za"], g[k + "_zb"], g[k + "_js"]
        cfg = ope.McConfig(n_paths=2 * mk.REPLAY_PAIRS, n_steps=mk.REPLAY_STEPS, seed=mk.SEED, scheme=scheme)
        payoff = ope.Payoff(kind, mk.K, 0, mk.BARRIER if kind == 2 else 0.0, 0.0)
        jumps = (lam, mk.JUMP[1], mk.JUMP[2]) if lam is not None else None
        run = ope.mc_price_ex(mk.S0, mk.T, mk.R, mk.Q, ope.HestonParams(**mk.HP), payoff, cfg, jumps,
                              ope.McOptions(0, replay=(za, zb, js if lam else None), want_paths=True))
        st, want = run.path_states, g[k + "_state"]
        np.testing.assert_allclose(st[..., :2], want[..., :2], rtol=1e-12, atol=1e-300, err_msg=k)
        np.testing.assert_allclose(run.path_payoffs, g[k + "_payoff"], rtol=0, atol=1e-12 * mk.S0, err_msg=k)


@pytest.mark.gpu
def test_gpu_cf_matches_the_fixture(ope, mk):
    g = load("cf.npz")
    hp = ope.HestonParams(**mk.HP)
    bp = ope.BatesParams(hp, *mk.JUMP)
    for i, t in enumerate(g["expiries"]):
        chain = [ope.OptionContract(mk.S0, float(k), float(t), mk.R, mk.Q, 0.0, ope.OptionType.Call)
                 for k in g["strikes"]]
        np.testing.assert_allclose(ope.batch_heston(chain, hp), g["heston"][i], rtol=0, atol=1e-12 * mk.S0)
        np.testing.assert_allclose(ope.batch_bates(chain, bp), g["bates"][i], rtol=0, atol=1e-12 * mk.S0)


@pytest.mark.gpu
def test_gpu_philox_draws_match_the_fixture(ope, mk):
    g = load("philox.npz")
    pairs = g["pairs"]
    for i, lam in enumerate((0.0, 0.5, 30.0)):
        for lo, hi in ((0, 8), (8, 16)):   # two contiguous runs of pair indices
            za, zb, js, nj = ope.debug_draws(mk.SEED, int(pairs[lo]), hi - lo, 19, 1.0, lam, mk.JUMP[1],
                                             mk.JUMP[2], 0)
            assert np.array_equal(nj.astype(np.int64), g[f"l{i}_n"][:, lo:hi])
            np.testing.assert_allclose(za, g[f"l{i}_za"][:, lo:hi], rtol=0, atol=1e-14 * 8)
            np.testing.assert_allclose(zb, g[f"l{i}_zb"][:, lo:hi], rtol=0, atol=1e-14 * 8)
            np.testing.assert_allclose(js, g[f"l{i}_js"][:, lo:hi], rtol=0, atol=1e-14)
