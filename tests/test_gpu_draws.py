"""GPU generator parity: Philox4x32-10 bit-exact against the Random123
known-answer vectors and the numpy oracle; Philox-mode normals, Poisson
counts and jump sizes against the oracle's restatement of the stream."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_device_is_gfx950(ope):
    info = ope.device_info()
    print("device:", info, flush=True)
    assert info["arch"].startswith("gfx950")
    assert info["compute_units"] >= 200


def test_philox_known_answers_on_device(ope):
    cases = [
        ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
        ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
        ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
         (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
    ]
    for ctr, key, want in cases:
        # rows 0 and 1: the two arrangements of the xors (the device kernel
        # runs odd rows through the split one, even rows through the QE
        # loops')
        out = ope.debug_philox(key[0] | (key[1] << 32), np.array([ctr, ctr], dtype=np.uint32))
        assert tuple(int(x) for x in out[0]) == want
        assert tuple(int(x) for x in out[1]) == want


def test_philox_random_counters_match_oracle(ope, ora):
    rng = np.random.default_rng(0)
    ctr = rng.integers(0, 2**32, size=(5000, 4), dtype=np.uint64).astype(np.uint32)
    seed = 0x0123456789ABCDEF
    out = ope.debug_philox(seed, ctr)
    ref = ora.philox4x32_10(*[ctr[:, i].astype(np.uint64) for i in range(4)], seed & 0xFFFFFFFF, seed >> 32)
    for i in range(4):
        assert np.array_equal(out[:, i].astype(np.uint64), ref[i])


@pytest.mark.parametrize("lam", [0.0, 0.5, 30.0])
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("n_steps", [7, 19])  # the FP32 8-step blocks: a tail alone; two blocks and an odd tail
def test_philox_draws_match_oracle(ope, ora, lam, precision, n_steps):
    seed, pair0, n_pairs = 42, 2**32 - 150, 300   # pairs straddle the counter's high word
    za, zb, js, nj = ope.debug_draws(seed, pair0, n_pairs, n_steps, 1.0, lam, -0.1, 0.15, precision)
    oza, ozb, ojs, onn = ora.philox_draws(seed, np.arange(pair0, pair0 + n_pairs, dtype=np.uint64), n_steps,
                                          dt=1.0 / n_steps, lam=lam, mu_j=-0.1, sigma_j=0.15)
    assert np.array_equal(nj.astype(np.int64), onn)
    if lam == 30.0:
        assert onn.max() >= 3  # the inverse-CDF tail walk is exercised
    if precision == 0:
        np.testing.assert_allclose(za, oza, rtol=0, atol=1e-14 * (1 + np.abs(oza).max()))
        np.testing.assert_allclose(zb, ozb, rtol=0, atol=1e-14 * (1 + np.abs(ozb).max()))
        np.testing.assert_allclose(js, ojs, rtol=0, atol=1e-14)
    else:
        # hardware FP32 log/sqrt/sin/cos: a few ulp of the FP64 values
        assert np.all(np.abs(za - oza) <= 5e-6 * (1 + np.abs(oza)))
        assert np.all(np.abs(zb - ozb) <= 5e-6 * (1 + np.abs(ozb)))
        np.testing.assert_allclose(js, ojs, rtol=0, atol=1e-6)
