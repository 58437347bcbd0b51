// Device-side path simulation for the Heston/Bates Monte Carlo engine.
//
// Included by exactly two translation units, each compiling its own copy
// (everything here has internal linkage):
//   mc_kernels.hip  Philox-mode pricing, built with -ffp-contract=on
//   mc_replay.hip   replay mode, built with -ffp-contract=off so the FP64 path
//                   repeats the reference's unfused arithmetic exactly
//
// Layout of the work: one lane owns one antithetic pair (mc.rs:177-193) and
// runs both siblings through all steps (mc.rs:243-273) with the path state
// in registers; strikes that share the simulation are priced from the
// terminal state; partial sums go to fixed slots per (tile, strike) in FP64.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mc_desc.h"

namespace ope {
namespace {

// ------------------------------------------------------------------ Philox
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11): 10 rounds of two 32x32->64 multiplies with a
// Weyl-sequence key schedule.  The key (seed) is uniform across the wave, so
// the key schedule lives in scalar registers.
constexpr uint32_t kPhiloxM0 = 0xD2511F53u, kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u, kPhiloxW1 = 0xBB67AE85u;
constexpr uint32_t kTagHeston = 0u, kTagBates = 1u;

__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)kPhiloxM0 * c.x;
        const uint64_t p1 = (uint64_t)kPhiloxM1 * c.z;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
        k0 += kPhiloxW0;
        k1 += kPhiloxW1;
    }
    return c;
}

// ------------------------------------------------------ FP32 hardware math
__device__ __forceinline__ float hw_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float hw_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float hw_log2(float x) { return __builtin_amdgcn_logf(x); }
__device__ __forceinline__ float hw_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float hw_sin_rev(float x) { return __builtin_amdgcn_sinf(x); }  // sin(2 pi x)
__device__ __forceinline__ float hw_cos_rev(float x) { return __builtin_amdgcn_cosf(x); }  // cos(2 pi x)

constexpr double kLog2e = 1.4426950408889634;
constexpr double kLn2 = 0.6931471805599453;

// --------------------------------------------------------- normal draws
// Box-Muller on two 32-bit words: u1 = (w0 + 1/2) 2^-32 in (0,1), u2 = w1 2^-32
// in [0,1); za = r cos(2 pi u2), zb = r sin(2 pi u2), r = sqrt(-2 ln u1).
// Both precisions consume the same words, so FP32 and FP64 runs see the
// same normals up to rounding.
__device__ __forceinline__ void box_muller_f64(uint32_t w0, uint32_t w1, double& za, double& zb) {
    const double u1 = ((double)w0 + 0.5) * 0x1p-32;
    const double u2 = (double)w1 * 0x1p-32;
    const double r = sqrt(-2.0 * log(u1));
    double s, c;
    sincospi(2.0 * u2, &s, &c);
    za = r * c;
    zb = r * s;
}

__device__ __forceinline__ void box_muller_f32(uint32_t w0, uint32_t w1, float& za, float& zb) {
    const float u1 = fmaf((float)w0, 0x1p-32f, 0x1p-33f);
    const float u2 = (float)w1 * 0x1p-32f;
    const float r = hw_sqrt(-1.38629436111989061f * hw_log2(u1));  // -2 ln2 log2(u1)
    za = r * hw_cos_rev(u2);
    zb = r * hw_sin_rev(u2);
}

// Inverse standard normal CDF, Wichura (1988) algorithm AS 241 (PPND16),
// about 1e-16 relative.  Used only for jump sizes (a rare branch).
__device__ __noinline__ double inv_normal_cdf(double p) {
    const double q = p - 0.5;
    if (fabs(q) <= 0.425) {
        const double r = 0.180625 - q * q;
        const double num = (((((((2.5090809287301226727e+3 * r + 3.3430575583588128105e+4) * r +
                                 6.7265770927008700853e+4) * r + 4.5921953931549871457e+4) * r +
                               1.3731693765509461125e+4) * r + 1.9715909503065514427e+3) * r +
                             1.3314166789178437745e+2) * r + 3.3871328727963666080e+0);
        const double den = (((((((5.2264952788528545610e+3 * r + 2.8729085735721942674e+4) * r +
                                 3.9307895800092710610e+4) * r + 2.1213794301586595867e+4) * r +
                               5.3941960214247511077e+3) * r + 6.8718700749205790830e+2) * r +
                             4.2313330701600911252e+1) * r + 1.0);
        return q * num / den;
    }
    double r = q < 0.0 ? p : 1.0 - p;
    r = sqrt(-log(r));
    double val;
    if (r <= 5.0) {
        r -= 1.6;
        const double num = (((((((7.74545014278341407640e-4 * r + 2.27238449892691845833e-2) * r +
                                 2.41780725177450611770e-1) * r + 1.27045825245236838258e+0) * r +
                               3.64784832476320460504e+0) * r + 5.76949722146069140550e+0) * r +
                             4.63033784615654529590e+0) * r + 1.42343711074968357734e+0);
        const double den = (((((((1.05075007164441684324e-9 * r + 5.47593808499534494600e-4) * r +
                                 1.51986665636164571966e-2) * r + 1.48103976427480074590e-1) * r +
                               6.89767334985100004550e-1) * r + 1.67638483018380384940e+0) * r +
                             2.05319162663775882187e+0) * r + 1.0);
        val = num / den;
    } else {
        r -= 5.0;
        const double num = (((((((2.01033439929228813265e-7 * r + 2.71155556874348757815e-5) * r +
                                 1.24266094738807843860e-3) * r + 2.65321895265761230930e-2) * r +
                               2.96560571828504891230e-1) * r + 1.78482653991729133580e+0) * r +
                             5.46378491116411436990e+0) * r + 6.65790464350110377720e+0);
        const double den = (((((((2.04426310338993978564e-15 * r + 1.42151175831644588870e-7) * r +
                                 1.84631831751005468180e-5) * r + 7.86869131145613259100e-4) * r +
                               1.48753612908506148525e-2) * r + 1.36929880922735805310e-1) * r +
                             5.99832206555887937690e-1) * r + 1.0);
        val = num / den;
    }
    return q < 0.0 ? -val : val;
}

// Exact Poisson count by inverse CDF of (w + 1/2) 2^-32.  The common n = 0
// case is decided by one integer compare against a host-computed threshold;
// the tail walks the CDF in FP64.
__device__ __noinline__ uint32_t poisson_tail(uint32_t w, double mu, double e0) {
    const double u = ((double)w + 0.5) * 0x1p-32;
    uint32_t k = 0;
    double p = e0, F = e0;
    while (u >= F && k < kPoissonKMax) {
        k += 1;
        p = p * mu / (double)k;
        F = F + p;
    }
    return k;
}

// Bates jump of one pair-step from Philox words 2 (count) and 3 (size):
// jump_sum = mu_j n + sigma_j sqrt(n) z (mc.rs:139-140), FP64 in both modes.
__device__ __forceinline__ double bates_jump(const McDesc& d, uint32_t w2, uint32_t w3, uint32_t& n) {
    n = 0;
    if (w2 >= d.pois_thr0) n = poisson_tail(w2, d.pois_mu, d.pois_e0);
    double j = 0.0;
    if (n != 0) {
        const double z = inv_normal_cdf(((double)w3 + 0.5) * 0x1p-32);
        j = d.mu_j * (double)n + d.sigma_j * sqrt((double)n) * z;
    }
    return j;
}

// Philox substream of one pair: key = seed, counter = (call, pair lo, pair hi, tag).
struct PairStream {
    uint32_t k0, k1, plo, phi;
    __device__ __forceinline__ uint4 call(uint32_t c0, uint32_t tag) const {
        return philox4x32_10(make_uint4(c0, plo, phi, tag), k0, k1);
    }
};

__device__ __forceinline__ PairStream make_stream(const McDesc& d, uint64_t pair) {
    PairStream s;
    s.k0 = (uint32_t)d.seed;
    s.k1 = (uint32_t)(d.seed >> 32);
    s.plo = (uint32_t)pair;
    s.phi = (uint32_t)(pair >> 32);
    return s;
}

// -------------------------------------------------- FP64 path (reference)
// Full-truncation Euler for one sibling (mc.rs:248-255).  z2 is the already
// rho-mixed normal of this sibling; negating both za and zb negates z2
// exactly, so the pair computes it once.
template <bool JUMPS>
__device__ __forceinline__ void euler_f64(const McDesc& d, double& ln_s, double& v, double z1, double z2,
                                          double jump) {
    const double v_pos = fmax(v, 0.0);
    const double sv = sqrt(v_pos);
    double inc = (d.drift - 0.5 * v_pos) * d.dt + sv * d.sqdt * z1;
    if (JUMPS) inc = inc + jump;
    ln_s = ln_s + inc;
    v = fmax(v + d.kappa * (d.theta - v_pos) * d.dt + d.sigma * sv * d.sqdt * z2, 0.0);
}

// Andersen (2008) QE variance step (mc.rs:288-309), psi_c = 1.5.
__device__ __forceinline__ double qe_variance_f64(const McDesc& d, double v_t, double z) {
    const double m = d.theta + (v_t - d.theta) * d.ekt;
    const double s2 = (v_t * d.sigma * d.sigma * d.ekt / d.kappa) * d.one_m_ekt + d.qe_c2;
    const double psi = fmax(s2 / fmax(m * m, 1e-300), 1e-12);
    if (psi <= 1.5) {
        const double inv_psi = 1.0 / psi;
        const double b2 = 2.0 * inv_psi - 1.0 + sqrt(fmax(2.0 * inv_psi * (2.0 * inv_psi - 1.0), 0.0));
        const double a = m / (1.0 + b2);
        const double b = sqrt(b2);
        const double bz = b + z;
        return a * (bz * bz);
    }
    const double p = (psi - 1.0) / (psi + 1.0);
    const double beta = (1.0 - p) / fmax(m, 1e-300);
    const double u = 0.5 * erfc(-z * 0.7071067811865476);  // ncdf, math.rs:16-19
    if (u <= p) return 0.0;
    return (1.0 / beta) * log((1.0 - p) / (1.0 - u));
}

// QE price update for one sibling (mc.rs:256-267): za drives v, zb drives S.
template <bool JUMPS>
__device__ __forceinline__ void qe_f64(const McDesc& d, double& ln_s, double& v, double za, double zb,
                                       double jump) {
    const double v_pos = fmax(v, 0.0);
    const double v_new = qe_variance_f64(d, v_pos, za);
    const double var_term = fmax(d.k3 * v_pos + d.k4 * v_new, 0.0);
    double inc = d.qe_a0 + d.k1 * v_pos + d.k2 * v_new + sqrt(var_term) * zb;
    if (JUMPS) inc = inc + jump;
    ln_s = ln_s + inc;
    v = v_new;
}

// ------------------------------------------------------- FP32-path math
// State y = log2(S/S0) kept in FP64 (one FP64 add per step removes the
// accumulation drift of a long FP32 sum); increments, variance and all
// transcendentals in FP32 on the hardware units.
template <bool JUMPS>
__device__ __forceinline__ void euler_f32(const McDesc& d, double& y, float& v, float z1, float z2,
                                          float jl) {
    const float vp = fmaxf(v, 0.0f);
    const float sv = hw_sqrt(vp);
    float dy = fmaf(d.eu_a2 * sv, z1, fmaf(d.eu_a1, vp, d.eu_a0));
    if (JUMPS) dy += jl;
    y += (double)dy;
    v = fmaxf(fmaf(d.eu_b2 * sv, z2, fmaf(d.eu_b1, vp, v + d.eu_b0)), 0.0f);
}

__device__ __forceinline__ float qe_variance_f32(const McDesc& d, float vp, float z) {
    const float m = fmaf(d.qe_m1, vp, d.qe_m0);
    const float s2 = fmaf(d.qe_s1, vp, d.qe_s0);
    const float psi = fmaxf(s2 * hw_rcp(fmaxf(m * m, 1e-36f)), 1e-12f);
    if (psi <= 1.5f) {
        const float t = 2.0f * hw_rcp(psi);
        const float b2 = (t - 1.0f) + hw_sqrt(fmaxf(t * (t - 1.0f), 0.0f));
        const float a = m * hw_rcp(1.0f + b2);
        const float bz = hw_sqrt(b2) + z;
        return a * (bz * bz);
    }
    // exponential branch: 1 - ncdf(z) = erfc(z/sqrt2)/2 directly (no cancellation)
    const float p = (psi - 1.0f) * hw_rcp(psi + 1.0f);
    const float q = 1.0f - p;
    const float one_m_u = 0.5f * erfcf(z * 0.70710678118654752f);
    if (one_m_u >= q) return 0.0f;
    return (hw_log2(q * hw_rcp(one_m_u)) * 0.69314718055994531f) * (fmaxf(m, 1e-36f) * hw_rcp(q));
}

template <bool JUMPS>
__device__ __forceinline__ void qe_f32(const McDesc& d, double& y, float& v, float za, float zb, float jl) {
    const float vp = fmaxf(v, 0.0f);
    const float vn = qe_variance_f32(d, vp, za);
    const float var = fmaxf(fmaf(d.qe_q3, vp, d.qe_q4 * vn), 0.0f);
    float dy = fmaf(hw_sqrt(var), zb, fmaf(d.qe_q2, vn, fmaf(d.qe_q1, vp, d.qe_q0)));
    if (JUMPS) dy += jl;
    y += (double)dy;
    v = vn;
}

// ------------------------------------------------------------ path state
struct PathState {
    double ln_s;   // ln S_T
    double v;      // v_T
    double sum;    // running_sum (Asian only, else NaN)
    double max;    // running_max (up-and-out only, else NaN)
    double x;      // payoff underlying: S_T, or running_sum / n_steps for Asian
    int ko;        // knocked out (up-and-out)
};

__device__ __forceinline__ double payoff_value(const McDesc& d, const PathState& e, double strike) {
    // disc * Payoff::eval (mc.rs:75-95, 275-277)
    const double intr = d.opt_put ? fmax(strike - e.x, 0.0) : fmax(e.x - strike, 0.0);
    return d.disc * (e.ko ? d.rebate : intr);
}

// Draw sources.  Philox: generated in registers.  Replay: read from arrays laid
// out [step][pair] (coalesced across the wave).
struct PhiloxSource {
    static constexpr bool kPhilox = true;
    PairStream st;
};

struct ReplaySource {
    static constexpr bool kPhilox = false;
    const double* za;
    const double* zb;
    const double* jump;  // may be null
    uint64_t n_pairs;
    uint64_t pair;
    __device__ __forceinline__ void get(uint32_t i, double& a, double& b, double& j) const {
        const uint64_t o = (uint64_t)i * n_pairs + pair;
        a = za[o];
        b = zb[o];
        j = jump ? jump[o] : 0.0;
    }
};

// --------------------------------------------- pair simulation, FP64 path
template <int SCHEME, bool JUMPS, int PAYOFF, class Src>
__device__ __forceinline__ void sim_pair_f64(const McDesc& d, const Src& src, bool anti, PathState& ea,
                                             PathState& eb) {
    double la = d.ln_spot, lb = d.ln_spot, va = d.v0, vb = d.v0;
    double suma = 0.0, sumb = 0.0;
    double mxa = -__builtin_huge_val(), mxb = -__builtin_huge_val();
    const uint32_t N = d.n_steps;

    auto step = [&](double za, double zb, double jump) {
        if (SCHEME == kSchemeEuler) {
            const double z2 = d.rho * za + d.sqrt_1m_rho2 * zb;
            euler_f64<JUMPS>(d, la, va, za, z2, jump);
            if (anti) euler_f64<JUMPS>(d, lb, vb, -za, -z2, jump);
        } else {
            qe_f64<JUMPS>(d, la, va, za, zb, jump);
            if (anti) qe_f64<JUMPS>(d, lb, vb, -za, -zb, jump);
        }
        if (PAYOFF == kPayAsian) {  // mc.rs:270-271
            suma += exp(la);
            if (anti) sumb += exp(lb);
        }
        if (PAYOFF == kPayUpAndOut) {  // running max of S, tracked as max of ln S (exp is monotone)
            mxa = fmax(mxa, la);
            if (anti) mxb = fmax(mxb, lb);
        }
    };

    if constexpr (Src::kPhilox) {
        if (JUMPS) {
            for (uint32_t i = 0; i < N; ++i) {
                const uint4 w = src.st.call(i, kTagBates);
                double za, zb;
                box_muller_f64(w.x, w.y, za, zb);
                uint32_t n;
                const double j = bates_jump(d, w.z, w.w, n);
                step(za, zb, j);
            }
        } else {
            uint32_t i = 0;
            for (; i + 1 < N; i += 2) {
                const uint4 w = src.st.call(i >> 1, kTagHeston);
                double za, zb;
                box_muller_f64(w.x, w.y, za, zb);
                step(za, zb, 0.0);
                box_muller_f64(w.z, w.w, za, zb);
                step(za, zb, 0.0);
            }
            if (i < N) {
                const uint4 w = src.st.call(i >> 1, kTagHeston);
                double za, zb;
                box_muller_f64(w.x, w.y, za, zb);
                step(za, zb, 0.0);
            }
        }
    } else {
        for (uint32_t i = 0; i < N; ++i) {
            double za, zb, j;
            src.get(i, za, zb, j);
            step(za, zb, j);
        }
    }

    const double nan = __builtin_nan("");
    auto finish = [&](PathState& e, double ln_s, double v, double sum, double mx) {
        e.ln_s = ln_s;
        e.v = v;
        e.sum = PAYOFF == kPayAsian ? sum : nan;
        e.max = nan;
        e.ko = 0;
        if (PAYOFF == kPayAsian) {
            e.x = sum / (double)N;
        } else {
            e.x = exp(ln_s);  // mc.rs:275
            if (PAYOFF == kPayUpAndOut) {
                const double running_max = fmax(d.spot, exp(mx));  // starts at spot (mc.rs:218)
                e.max = running_max;
                e.ko = running_max >= d.barrier;
            }
        }
    };
    finish(ea, la, va, suma, mxa);
    if (anti) finish(eb, lb, vb, sumb, mxb);
}

// --------------------------------------------- pair simulation, FP32 path
template <int SCHEME, bool JUMPS, int PAYOFF, class Src>
__device__ __forceinline__ void sim_pair_f32(const McDesc& d, const Src& src, bool anti, PathState& ea,
                                             PathState& eb) {
    double ya = 0.0, yb = 0.0;
    float va = d.f_v0, vb = d.f_v0;
    float suma = 0.0f, sumb = 0.0f;
    double mxa = -__builtin_huge_val(), mxb = -__builtin_huge_val();
    const uint32_t N = d.n_steps;

    auto step = [&](float za, float zb, float jl) {
        if (SCHEME == kSchemeEuler) {
            const float z2 = fmaf(d.f_rho, za, d.f_rhoc * zb);
            euler_f32<JUMPS>(d, ya, va, za, z2, jl);
            if (anti) euler_f32<JUMPS>(d, yb, vb, -za, -z2, jl);
        } else {
            qe_f32<JUMPS>(d, ya, va, za, zb, jl);
            if (anti) qe_f32<JUMPS>(d, yb, vb, -za, -zb, jl);
        }
        if (PAYOFF == kPayAsian) {
            suma += hw_exp2((float)ya);
            if (anti) sumb += hw_exp2((float)yb);
        }
        if (PAYOFF == kPayUpAndOut) {
            mxa = fmax(mxa, ya);
            if (anti) mxb = fmax(mxb, yb);
        }
    };

    if constexpr (Src::kPhilox) {
        if (JUMPS) {
            for (uint32_t i = 0; i < N; ++i) {
                const uint4 w = src.st.call(i, kTagBates);
                float za, zb;
                box_muller_f32(w.x, w.y, za, zb);
                uint32_t n;
                const double j = bates_jump(d, w.z, w.w, n);
                step(za, zb, (float)(j * kLog2e));
            }
        } else {
            uint32_t i = 0;
            for (; i + 1 < N; i += 2) {
                const uint4 w = src.st.call(i >> 1, kTagHeston);
                float za, zb;
                box_muller_f32(w.x, w.y, za, zb);
                step(za, zb, 0.0f);
                box_muller_f32(w.z, w.w, za, zb);
                step(za, zb, 0.0f);
            }
            if (i < N) {
                const uint4 w = src.st.call(i >> 1, kTagHeston);
                float za, zb;
                box_muller_f32(w.x, w.y, za, zb);
                step(za, zb, 0.0f);
            }
        }
    } else {
        for (uint32_t i = 0; i < N; ++i) {
            double za, zb, j;
            src.get(i, za, zb, j);
            step((float)za, (float)zb, (float)(j * kLog2e));
        }
    }

    const double nan = __builtin_nan("");
    auto finish = [&](PathState& e, double y, float v, float sum, double mx) {
        e.ln_s = d.ln_spot + y * kLn2;
        e.v = (double)v;
        e.sum = PAYOFF == kPayAsian ? d.spot * (double)sum : nan;
        e.max = nan;
        e.ko = 0;
        if (PAYOFF == kPayAsian) {
            e.x = d.spot * (double)sum / (double)N;
        } else {
            e.x = d.spot * exp2(y);
            if (PAYOFF == kPayUpAndOut) {
                // barrier test in FP64 on the FP64 log-price maximum
                e.ko = (d.spot >= d.barrier) || (mx >= d.lb2_barrier);
                e.max = fmax(d.spot, d.spot * exp2(mx));
            }
        }
    };
    finish(ea, ya, va, suma, mxa);
    if (anti) finish(eb, yb, vb, sumb, mxb);
}

// ------------------------------------------------------------ reductions
__device__ __forceinline__ void wave_sum2(double& s, double& q) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        s += __shfl_xor(s, m, 64);
        q += __shfl_xor(q, m, 64);
    }
}

// Per-path outputs (replay mode only).
struct PathSink {
    double* payoff;  // [pair][2]
    double* state;   // [pair][2][4]
};

// One tile: kBlock * PPT pairs of one request.  Thread t simulates pairs
// tile_base + j*kBlock + t (j < PPT), then for every strike the block sums
// the discounted payoffs and their squares in a fixed order (per-thread j
// order, xor-butterfly over the wave, waves 0..3) and writes one FP64
// (sum, sum_sq) slot.  The order depends on the request alone, never on
// what else shares the launch.
template <class Real, int SCHEME, bool JUMPS, int PAYOFF, int PPT, class MakeSrc>
__device__ __forceinline__ void price_tile(const McDesc& d, uint32_t tile, const double* __restrict__ strikes,
                                           double2* __restrict__ slots, const MakeSrc& make_src,
                                           const PathSink& sink, double2* lds) {
    const int tid = threadIdx.x;
    const bool anti = d.antithetic != 0;
    const uint64_t base = (uint64_t)tile * (uint64_t)(kBlock * PPT);
    PathState ea[PPT], eb[PPT];
    bool valid[PPT];
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        const uint64_t pair = base + (uint64_t)j * kBlock + (uint64_t)tid;
        valid[j] = pair < d.n_pairs;
        if (valid[j]) {
            const auto src = make_src(d, pair);
            if constexpr (sizeof(Real) == 8)
                sim_pair_f64<SCHEME, JUMPS, PAYOFF>(d, src, anti, ea[j], eb[j]);
            else
                sim_pair_f32<SCHEME, JUMPS, PAYOFF>(d, src, anti, ea[j], eb[j]);
            if (sink.payoff) {
                const double k0 = strikes[d.strike_off];
                sink.payoff[pair * 2] = payoff_value(d, ea[j], k0);
                sink.payoff[pair * 2 + 1] = anti ? payoff_value(d, eb[j], k0) : __builtin_nan("");
            }
            if (sink.state) {
                double* o = sink.state + pair * 8;
                o[0] = ea[j].ln_s; o[1] = ea[j].v; o[2] = ea[j].sum; o[3] = ea[j].max;
                if (anti) { o[4] = eb[j].ln_s; o[5] = eb[j].v; o[6] = eb[j].sum; o[7] = eb[j].max; }
            }
        }
    }

    const int lane = tid & 63, wave = tid >> 6;
    const uint32_t ns = d.n_strikes;
    for (uint32_t k = 0; k < ns; ++k) {
        const double K = strikes[d.strike_off + k];
        double s = 0.0, q = 0.0;
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            if (valid[j]) {
                const double pa = payoff_value(d, ea[j], K);
                s += pa;
                q += pa * pa;
                if (anti) {
                    const double pb = payoff_value(d, eb[j], K);
                    s += pb;
                    q += pb * pb;
                }
            }
        }
        wave_sum2(s, q);
        if (lane == 0) lds[k * 4 + wave] = make_double2(s, q);
    }
    __syncthreads();
    for (uint32_t k = tid; k < ns; k += kBlock) {
        const double2 a = lds[k * 4 + 0], b = lds[k * 4 + 1], c = lds[k * 4 + 2], e = lds[k * 4 + 3];
        slots[d.slot_off + (uint64_t)tile * ns + k] =
            make_double2(((a.x + b.x) + c.x) + e.x, ((a.y + b.y) + c.y) + e.y);
    }
}

}  // namespace
}  // namespace ope
