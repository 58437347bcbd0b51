"""Algorithmic work of the reference's Monte Carlo step body and analytic pricer.

Counting rule: the floating-point operations of mc.rs's step body after loop
invariants are hoisted; add / sub / mul / div / min / max / compare = 1,
a*b+c = 2, each transcendental (exp, ln, sqrt, sin, cos, erfc) = 1.  Draw
work per antithetic pair (Box-Muller, mc.rs:312-318; Knuth's Poisson test,
mc.rs:326-329) is split over the two siblings.  Integer generator work
(Philox) is not counted.

Per path-step:
  Euler step (mc.rs:245-255, 270-272)                  25 flops
  QE step, quadratic branch (mc.rs:256-267, 288-302)   40 flops
  Box-Muller per pair (12) / 2                          6 flops
  Poisson test per pair (3) / 2, Bates only             1.5 flops
giving W = 31 (Euler, Heston), 32.5 (Euler, Bates), 46 (QE, Heston),
47.5 (QE, Bates).  Per path at expiry: 1 exp, plus 6 flops per strike
(intrinsic, discount, sum, sum of squares), plus 1 for the Asian mean or the
barrier select.

Analytic pricer (heston.rs:82-229, bates.rs:49-83), per integrand
evaluation, same counting rule, complex product = 6, complex quotient = 11
(num-complex's naive forms), complex exp = 5 (exp, cos, sin, 2 mul), complex
ln = 3 (hypot, ln, atan2):
  map u = (1-t)/t and its guards 7; stable_cf 150 (xi 11, d with the stable
  square root 27, g 15, e^{-dt} 9, A 19, C 16, D 34, r phi i t 10, final
  exp 9); P1 normalisation 11 on half the evaluations; e^{iux} cf 12;
  division by iu 11
giving W_CF = 185.5 (Heston), + 57 for the jump factor and its product =
242.5 (Bates).  An integral with p panels evaluated 15 (2p - 1) nodes.

MI355X vector peaks (datasheet): FP32 157.3 TFLOP/s (256 CUs x 4 SIMDs x 32
FMA lanes per clock x 2 x 2.4 GHz), FP64 78.6 TFLOP/s; HBM3E 8 TB/s.
"""

W_STEP = {("euler", False): 31.0, ("euler", True): 32.5, ("qe", False): 46.0, ("qe", True): 47.5}
PEAK_TFLOPS = {"f32": 157.3, "f64": 78.6}
PEAK_HBM_GBS = 8000.0


def flops_per_path(n_steps: int, scheme: str, jumps: bool, n_strikes: int = 1, payoff: str = "european") -> float:
    extra = 1.0 if payoff in ("asian", "up_and_out") else 0.0
    return n_steps * W_STEP[(scheme, jumps)] + 1.0 + 6.0 * n_strikes + extra


def flops(n_paths: int, n_steps: int, scheme: str, jumps: bool, n_strikes: int = 1, payoff: str = "european") -> float:
    """Algorithmic flops of one simulation of n_paths paths (both antithetic siblings counted)."""
    return n_paths * flops_per_path(n_steps, scheme, jumps, n_strikes, payoff)


W_CF = {"heston": 185.5, "bates": 242.5}


def cf_flops(model: str, panels_p1: int, panels_p2: int) -> float:
    """Algorithmic flops of one analytic price from its quadrature panel counts."""
    return W_CF[model] * 15.0 * ((2 * panels_p1 - 1) + (2 * panels_p2 - 1))
