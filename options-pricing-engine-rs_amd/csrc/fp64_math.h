// FP64 elementary functions for the Philox-mode FP64 path kernels, for the
// argument ranges that path produces.  Each is a few FMAs on top of a
// hardware approximation or an exact reduction, instead of the general
// device-library function (whose log and sincospi carry double-double
// compensation and general range reduction).  Accuracy, checked against the
// mpmath on the GPU (ope_debug_fp64_math): within 2 ulp, the square root
// correctly rounded on every tested input, the normal tail within 8 ulp.
//
// Methods:
//   rcp_f64     v_rcp_f64 refined by two Newton-Raphson steps, y += y (1 - x y).
//   rsqrt_f64   v_rsq_f64 refined by two Newton-Raphson steps,
//               y += (y/2)(1 - x y^2).
//   sqrt_f64    for x = 0 or a positive normal x: from y0 = v_rsq_f64(x),
//               the coupled iteration for g -> sqrt(x) and h -> 1/(2 sqrt(x))
//               (Goldschmidt; one step r = 1/2 - g h, g += g r, h += h r), then
//               two corrections g += (x - g^2) h.  No scaling for denormal or
//               infinite input; x = 0 gives 0 (the rsq input is floored at
//               the least normal, so g = 0 y0 = 0).
//   log_f64     x = 2^k m with m in [sqrt(1/2), sqrt 2) (frexp, exact); m
//               rounded to i/128 picks a table entry (gen_fp64_tables.py)
//               with inv ~ 128/i and -ln(inv) in two parts; ln m =
//               ln1p(r) - ln(inv) with r = m inv - 1 (one fma, |r| <= 0.0056,
//               exact where inv = 1, so ln(1 + eps) keeps its relative
//               accuracy), ln1p(r) = r + r^2 q(r) with the Taylor series to
//               degree 7, and k ln2 added as k ln2_hi + k ln2_lo (Cody and
//               Waite; Tang's table-driven method).
//   exp_f64     x = (64 e + j) ln2/64 + r with k = 64 e + j = rint(64 x/ln2),
//               the reduction in two parts (k ln2/64_hi exact), e^r - 1 to
//               degree 6 (|r| <= ln2/128), times 2^(j/64) from a table in
//               two parts, scaled by 2^e (ldexp; Tang's method).
//   ncdf_neg_abs_f64  Phi(-|z|) for |z| <= 6.8 (every Box-Muller normal of
//               the engine: u1 >= 2^-33): exp(-z^2/2) t g(t) with t = 1/(1 +
//               |z|/5) and g a degree-16 polynomial in the affine image of t
//               over [-1, 1] (Chebyshev fit of Phi(-y) e^{y^2/2}/t by mpmath,
//               error below 1e-17), z^2/2 split into a rounded part and its
//               fma remainder; within 8 ulp (7e-16 relative, mostly the
//               roundings of t and of the exponential).
//   sincos_rev23  sin and cos of 2 pi k 2^-23 for a 23-bit integer k: the
//               nearest 1/256 turn is taken in integer arithmetic (exact)
//               and looked up (gen_fp64_tables.py), the remaining angle (at
//               most pi/256) goes through Taylor polynomials to degree 7
//               (sine) and 6 (cosine), and the angle-sum formulas combine
//               the two.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fp64_tables.h"

namespace ope {
namespace {

constexpr double kLn2Hi = 0.6931471806019545;       // ln 2 to 29 significant bits: k ln2_hi is exact
constexpr double kLn2Lo = -4.2009150726810846e-11;  // ln 2 - ln2_hi

__device__ __forceinline__ double rcp_f64(double x) {
    double y = __builtin_amdgcn_rcp(x);
    double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
    e = fma(-x, y, 1.0);
    return fma(y, e, y);
}

// 1/sqrt(x) for a positive normal x.
__device__ __forceinline__ double rsqrt_f64(double x) {
    double y = __builtin_amdgcn_rsq(x);
    double e = fma(-x * y, y, 1.0);
    y = fma(0.5 * y, e, y);
    e = fma(-x * y, y, 1.0);
    return fma(0.5 * y, e, y);
}

// sqrt(x) for x = 0 or a positive normal x.
__device__ __forceinline__ double sqrt_f64(double x) {
    const double y0 = __builtin_amdgcn_rsq(fmax(x, 0x1p-1022));
    double g = x * y0, h = 0.5 * y0;
    const double r = fma(-g, h, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    g = fma(fma(-g, g, x), h, g);
    return fma(fma(-g, g, x), h, g);
}

// Natural logarithm of a positive normal x.
__device__ __forceinline__ double log_f64(double x) {
    double m = __builtin_amdgcn_frexp_mant(x);  // [1/2, 1)
    int k = __builtin_amdgcn_frexp_exp(x);
    const bool lo = m < 0.70710678118654752;
    m = lo ? m + m : m;  // [sqrt(1/2), sqrt 2)
    k = lo ? k - 1 : k;
    const int i = (int)__builtin_rint(m * 128.0);  // 91..181
    const LogEntry t = kLogTab[i - 91];
    const double r = fma(m, t.inv, -1.0);  // |r| <= 0.0056; exact at i = 128
    double q = 0.14285714285714285;        // ln1p(r) = r + r^2 q(r), q to degree 5
    q = fma(q, r, -0.16666666666666666);
    q = fma(q, r, 0.2);
    q = fma(q, r, -0.25);
    q = fma(q, r, 0.33333333333333331);
    q = fma(q, r, -0.5);
    const double kd = (double)k;
    const double tail = fma(r * r, q, fma(kd, kLn2Lo, t.llo));
    return fma(kd, kLn2Hi, t.lhi + (r + tail));
}

// e^x for |x| < 700: x = (64 e + j) ln2/64 + r, |r| <= ln2/128, e^x =
// 2^e 2^(j/64) e^r with e^r - 1 to degree 6.
__device__ __forceinline__ double exp_f64(double x) {
    const double kd = __builtin_rint(x * 92.332482616893658);  // 64 / ln 2
    const int k = (int)kd;
    const double r = fma(-kd, kLn2By64Lo, fma(-kd, kLn2By64Hi, x));
    double p = 1.3888888888888889e-03;  // 1/720
    p = fma(p, r, 8.3333333333333333e-03);
    p = fma(p, r, 4.1666666666666667e-02);
    p = fma(p, r, 1.6666666666666667e-01);
    p = fma(p, r, 0.5);
    p = fma(r * r, p, r);  // e^r - 1
    const double2 t = kExp2Tab[k & 63];
    return __builtin_amdgcn_ldexp(t.x + fma(t.x, p, t.y), k >> 6);
}

// Phi(-|z|) = erfc(|z|/sqrt 2)/2 for |z| <= 6.8.
__device__ __forceinline__ double ncdf_neg_abs_f64(double z) {
    const double y = fabs(z);
    const double t = rcp_f64(fma(0.2, y, 1.0));
    const double x = fma(3.4705882352941178, t, -2.4705882352941178);  // t in [1/2.36, 1] -> [-1, 1]
    double g = -8.7157183142503107e-12;
    g = fma(g, x, 1.6224162689825668e-11);
    g = fma(g, x, 2.3385148510312122e-10);
    g = fma(g, x, -6.264983711853694e-10);
    g = fma(g, x, -5.9473821121120224e-09);
    g = fma(g, x, 1.6663398724031173e-08);
    g = fma(g, x, 1.8071965454105019e-07);
    g = fma(g, x, -2.4348625623622599e-07);
    g = fma(g, x, -6.257901227652141e-06);
    g = fma(g, x, -1.318750259107785e-05);
    g = fma(g, x, 0.00016229816314654387);
    g = fma(g, x, 0.0016220572450152027);
    g = fma(g, x, 0.0082198281709849313);
    g = fma(g, x, 0.028762000825588444);
    g = fma(g, x, 0.075381249934953001);
    g = fma(g, x, 0.15181031902153735);
    g = fma(g, x, 0.23406174447831679);
    const double h = -0.5 * y;
    const double e = y * h, elo = fma(y, h, -e);  // -y^2/2 = e + elo exactly
    const double ex = exp_f64(e);
    return (t * g) * fma(ex, elo, ex);
}

// sin and cos of 2 pi k 2^-23, k in [0, 2^23): the nearest 1/256 turn j
// (integer arithmetic, exact) from kSinCos256, the rest b = 2 pi l 2^-23,
// |b| <= pi/256, through sin b (degree 7) and cos b - 1 (degree 6), and the
// angle-sum formulas with the cosine's deviation from 1 kept separate, so
// that a zero of the table stays a zero (exact zeros of the sine at half
// turns and of the cosine at odd quarter turns).
__device__ __forceinline__ void sincos_rev23(uint32_t k, double& sn, double& cs) {
    const uint32_t j = (k + (1u << 14)) >> 15;              // 0..256
    const int32_t l = (int32_t)k - (int32_t)(j << 15);      // [-2^14, 2^14)
    const double2 t = kSinCos256[j & 255u];                 // (sin, cos) of 2 pi j / 256
    const double b = (double)l * 7.4901405658478570e-07;    // 2 pi 2^-23
    const double b2 = b * b;
    const double sb = fma(b * b2, fma(b2, fma(b2, -1.9841269841269841e-04, 8.3333333333333333e-03),
                                      -1.6666666666666667e-01), b);
    const double cm = b2 * fma(b2, fma(b2, -1.3888888888888889e-03, 4.1666666666666667e-02), -0.5);
    sn = fma(t.y, sb, fma(t.x, cm, t.x));
    cs = fma(-t.x, sb, fma(t.y, cm, t.y));
}

}  // namespace
}  // namespace ope
