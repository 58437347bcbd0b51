// Request descriptor shared by the host planner (capi.cpp) and the kernels.
//
// One McDesc per priced simulation.  The host computes every loop invariant
// once, in FP64, with the same expressions and evaluation order as the
// reference (mc.rs:153-157 for dt/sqdt/drift, mc.rs:230-238 for the QE
// K0..K4, mc.rs:291-294 for the QE moment terms), so that the FP64 kernels
// reproduce the reference's per-step arithmetic exactly.  The FP32-path
// constants are the same quantities folded into log2 space and rounded once.
#pragma once

#include <stdint.h>

namespace ope {

enum : uint32_t { kSchemeEuler = 0, kSchemeQE = 1 };
enum : uint32_t { kPayEuropean = 0, kPayAsian = 1, kPayUpAndOut = 2 };

constexpr int kBlock = 256;         // threads per tile block
constexpr int kMaxStrikes = 256;    // strikes priced from one simulation
constexpr uint32_t kPoissonKMax = 1024;

struct alignas(16) McDesc {
    // ---- integers
    uint64_t seed;        // Philox key
    uint64_t n_pairs;     // antithetic pairs (or single paths when !antithetic)
    uint64_t slot_off;    // first partial-sum slot of this request
    uint32_t n_steps;
    uint32_t n_tiles;     // tiles of kBlock * ppt pairs
    uint32_t ppt;         // pairs per thread: 1 or 4 (a function of n_pairs only)
    uint32_t scheme;      // kScheme*
    uint32_t payoff;      // kPay*
    uint32_t opt_put;     // 0 call, 1 put
    uint32_t antithetic;  // 0 / 1
    uint32_t jumps;       // 1 when lambda > 0 (Bates stream), else Heston stream
    uint32_t n_strikes;
    uint32_t strike_off;  // into the strike array
    uint32_t out_off;     // into the (sum, sum_sq) result array
    uint32_t pois_thr0;   // w < thr0  <=>  (w + 0.5) 2^-32 < e^-mu  (Poisson count 0)
    uint32_t pois_gate;   // thr0 >> 23: 9 high bits below it mean count 0 whatever the low bits
    uint32_t pois_thr1;   // thr0 <= w < thr1  =>  count 1 (the CDF walk's first step, done on the host)
    uint32_t flags;       // OPE_FLAG_* of the call (bit 0: FP64 barrier decision near B)

    // ---- FP64 reference-arithmetic constants
    double spot, ln_spot, disc, dt, sqdt, drift;
    double v0, kappa, theta, sigma, rho, sqrt_1m_rho2;
    double ekt, one_m_ekt, qe_c2;  // exp(-kappa dt), 1 - ekt, (theta s s / 2k)(1-ekt)^2
    double qe_a0;                  // drift*dt + K0
    double k1, k2, k3, k4;
    double barrier, rebate;
    double pois_mu, pois_e0, mu_j, sigma_j;
    double lb2_barrier;            // log2(barrier / spot), FP32-path barrier test

    // ---- FP64 Philox-mode constants (hoisted forms of the same quantities)
    double ed_a0, ed_a1;           // Euler ln-price: inc = a0 + a1 v + sqrt(v) sqrt(dt) z1
    double ed_b0, ed_c1, ed_sb;    // Euler variance: v' = max(b0 + c1 v + sqrt(v) sb z2, 0)
    double qd_m0, qd_s1, qd_s0;    // QE: m = ekt v + m0, s2 / 2 = s1 v + s0
    double qd_sq3;                 // QE: sqrt(max(K3, 0)) (K3 = K4)

    // ---- FP32-path constants (log2(S/S0) state)
    float f_v0;
    float eu_a0, eu_a1, eu_a2;     // dy = a0 + a1 v+ + a2 sqrt(v+) z1
    float eu_b0, eu_b1, eu_b2;     // v' = max(v + b0 + b1 v+ + b2 sqrt(v+) z2, 0)
    float eu_c1;                   // 1 + b1 (v' for v >= 0: max(b0 + c1 v + b2 sqrt(v) z2, 0))
    float eu_r2, eu_kb, eu_phi;    // -2 ln2 a2^2, b2/a2, acos(rho)/(2 pi): mixed Box-Muller
    float f_rho, f_rhoc;           // z2 = rho za + sqrt(1-rho^2) zb
    float qe_m1, qe_m0;            // m = m1 v+ + m0
    float qe_s1, qe_s0;            // s2 / 2 = s1 v+ + s0 (halved on the host)
    float qe_q0, qe_q1, qe_q2;     // dy = q0 + q1 v+ + q2 v' + sq3 sqrt(v+ + v') zb
    float qe_sq3;                  // sqrt(max(q3, 0)), q3 = q4 the log2-scaled K3 = K4
    float qe_vlow;                 // QE regrouping: v below it has u = psi/2 > 0.1 (0: no such v)
    float f_mu_j2, f_sigma_j2;     // log2e * mu_j, log2e * sigma_j
};

}  // namespace ope
