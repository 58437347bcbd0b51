// Analytic Heston / Bates European prices on the GPU (heston.rs:61-229,
// bates.rs:29-83): Gil-Pelaez inversion of the Albrecher-form characteristic
// function with the reference's globally adaptive Gauss-Kronrod 15/7
// quadrature on the u = (1 - t)/t map of [0, inf) onto [0, 1].
//
// One wavefront (one 64-thread workgroup) per option.  Lanes 0-31 run the P1
// integral, lanes 32-63 the P2 integral.  In each bisection step, lanes 0-29
// of a half evaluate the 15 nodes of both children of the panel being split;
// one lane per child then forms the K15 and G7 sums in the reference's order.
// The panel list lives in LDS and is kept in the reference's order
// (swap_remove, then push left, push right), and the panel to split is the
// one with the largest error, the last one on ties (Iterator::max_by), found
// by a 32-lane butterfly.  So the device takes the same bisection decisions
// as the reference, and a price differs from the reference's only through
// the last bits of the FP64 exp / log / sin / cos / atan2 / hypot.
// Built with -ffp-contract=off (rustc never fuses).
#include <hip/hip_runtime.h>

#include "cf_desc.h"

namespace ope {
namespace {

// ---- complex arithmetic with num-complex 0.4 semantics: naive product and
// quotient, exp via (e^re cos im, e^re sin im), ln via (ln hypot, atan2).
struct cpx {
    double re, im;
};
__device__ __forceinline__ cpx C(double re, double im) { return cpx{re, im}; }
__device__ __forceinline__ cpx cadd(cpx a, cpx b) { return C(a.re + b.re, a.im + b.im); }
__device__ __forceinline__ cpx csub(cpx a, cpx b) { return C(a.re - b.re, a.im - b.im); }
__device__ __forceinline__ cpx cmul(cpx a, cpx b) { return C(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
__device__ __forceinline__ cpx cdiv(cpx a, cpx b) {
    const double n = b.re * b.re + b.im * b.im;
    const double re = a.re * b.re + a.im * b.im;
    const double im = a.im * b.re - a.re * b.im;
    return C(re / n, im / n);
}
__device__ __forceinline__ cpx cscale(cpx a, double s) { return C(a.re * s, a.im * s); }  // Complex * f64
__device__ __forceinline__ cpx cradd(cpx a, double s) { return C(a.re + s, a.im); }       // Complex + f64
__device__ __forceinline__ cpx crsub(double s, cpx a) { return C(s - a.re, -a.im); }      // f64 - Complex
__device__ __forceinline__ cpx cexp_(cpx z) {
    const double e = exp(z.re);
    double s, c;
    sincos(z.im, &s, &c);  // one argument reduction for both
    return C(e * c, e * s);
}
__device__ __forceinline__ cpx clog_(cpx z) { return C(log(hypot(z.re, z.im)), atan2(z.im, z.re)); }

// heston.rs:155-175: the square root that computes the cancellation-free
// component first and the other one from re*im = b/2.
__device__ cpx fast_csqrt(cpx z) {
    const double a = z.re, b = z.im;
    if (b == 0.0) {
        if (a >= 0.0) return C(sqrt(a), b);
        const double m = sqrt(-a);
        return C(0.0, signbit(b) ? -m : m);
    }
    const double r = hypot(a, b);
    if (a >= 0.0) {
        const double re_out = sqrt((r + a) / 2.0);
        return C(re_out, b / (2.0 * re_out));
    }
    const double im_out = copysign(sqrt((r - a) / 2.0), b);
    return C(b / (2.0 * im_out), im_out);
}

// heston.rs:113-127, the Albrecher et al. (2007) form.
__device__ cpx stable_cf(cpx phi, const CfDesc& p) {
    const cpx I = C(0.0, 1.0);
    const double t = p.expiry;
    const cpx xi = crsub(p.kappa, cmul(cscale(phi, p.rho * p.sigma), I));
    const cpx d = fast_csqrt(cadd(cmul(xi, xi), cmul(cscale(phi, p.sigma * p.sigma), cadd(phi, I))));
    const cpx g = cdiv(csub(xi, d), cadd(xi, d));
    const cpx edt = cexp_(cscale(C(-d.re, -d.im), t));
    const cpx a = cdiv(cradd(cmul(g, edt), -1.0), cradd(g, -1.0));
    const cpx c = cscale(csub(cscale(csub(xi, d), t), cscale(clog_(a), 2.0)), p.kappa * p.theta / (p.sigma * p.sigma));
    const cpx dd = cdiv(cmul(cscale(csub(xi, d), p.v0), crsub(1.0, edt)),
                        cscale(crsub(1.0, cmul(g, edt)), p.sigma * p.sigma));
    const cpx rpi = cscale(cmul(cscale(phi, p.rate), I), t);
    return cexp_(cadd(cadd(rpi, c), dd));
}

// bates.rs:77-83: Merton log-normal jump factor.
__device__ cpx jump_cf(cpx phi, const CfDesc& p) {
    const cpx I = C(0.0, 1.0);
    const double comp = exp(p.mu_j + 0.5 * p.sigma_j * p.sigma_j) - 1.0;
    const cpx jump = cexp_(csub(cscale(cmul(phi, I), p.mu_j),
                                cscale(cscale(cmul(cscale(phi, 0.5), phi), p.sigma_j), p.sigma_j)));
    const cpx inner = csub(cradd(jump, -1.0), cscale(cmul(I, phi), comp));
    return cexp_(cscale(inner, p.lambda * p.expiry));
}

// d ln J(phi) / d(lambda, mu_j, sigma_j) of the jump factor above, J =
// exp(lambda T (E - 1 - i phi k)), E = exp(i phi mu_j - sigma_j^2 phi^2 / 2),
// k = e^(mu_j + sigma_j^2/2) - 1:
//   lambda:  T (E - 1 - i phi k)
//   mu_j:    lambda T i phi (E - e^(mu_j + sigma_j^2/2))
//   sigma_j: -lambda T sigma_j (phi^2 E + i phi e^(mu_j + sigma_j^2/2))
// The integrands' derivatives are Re(z g) for the integrand z: only J
// depends on the jump parameters, and the P1 normaliser's factor J(-i) is 1
// with zero derivative (E(-i) = e^(mu_j + sigma_j^2/2)).  These are the
// quantities the reference's forward-mode pass (ad.rs:218-234, 609-630)
// propagates through the same expressions.
__device__ void jump_dlog(cpx phi, const CfDesc& p, cpx g[3]) {
    const cpx I = C(0.0, 1.0);
    const double ec = exp(p.mu_j + 0.5 * p.sigma_j * p.sigma_j);
    const cpx E = cexp_(csub(cscale(cmul(phi, I), p.mu_j),
                             cscale(cscale(cmul(cscale(phi, 0.5), phi), p.sigma_j), p.sigma_j)));
    const cpx iphi = cmul(I, phi);
    const cpx inner = csub(cradd(E, -1.0), cscale(iphi, ec - 1.0));
    g[0] = cscale(inner, p.expiry);
    g[1] = cscale(cmul(iphi, cradd(E, -ec)), p.lambda * p.expiry);
    g[2] = cscale(cadd(cmul(cmul(phi, phi), E), cscale(iphi, ec)), -p.lambda * p.expiry * p.sigma_j);
}

__device__ __forceinline__ cpx model_cf(cpx phi, const CfDesc& p) {
    cpx cf = stable_cf(phi, p);
    if (p.bates) cf = cmul(cf, jump_cf(phi, p));
    return cf;
}

// heston.rs:101-109 / bates.rs:65-73, then the map of heston.rs:220-229.
// With DERIV (Bates), dv[k] receives the derivative of the mapped integrand
// with respect to (lambda, mu_j, sigma_j), 0 where it or the value is not
// finite (ad.rs:673-684); the value keeps the pricer's own guard, so the
// quadrature takes the same bisection decisions as the price alone.
template <bool DERIV>
__device__ double mapped(double t, const CfDesc& p, double x, bool is_p1, cpx norm, double* dv) {
    if (DERIV) dv[0] = dv[1] = dv[2] = 0.0;
    if (t <= 0.0) return 0.0;
    const double u = (1.0 - t) / t;
    if (u < 1e-12) return 0.0;
    const cpx phi = is_p1 ? C(u, -1.0) : C(u, 0.0);
    cpx cf = model_cf(phi, p);
    if (is_p1) cf = cdiv(cf, norm);
    const cpx num = cmul(cexp_(C(0.0, u * x)), cf);
    const cpx z = cdiv(num, C(0.0, u));
    const double v = z.re / (t * t);
    if (!isfinite(v)) return 0.0;
    if (DERIV) {
        cpx g[3];
        jump_dlog(phi, p, g);
        for (int k = 0; k < 3; ++k) {
            const double dk = cmul(z, g[k]).re / (t * t);
            dv[k] = isfinite(dk) ? dk : 0.0;
        }
    }
    return v;
}

// heston.rs:28-59
__constant__ double kGkNodes[15] = {
    0.0,
    0.2077849550078985, -0.2077849550078985,
    0.4058451513773972, -0.4058451513773972,
    0.5860872354676911, -0.5860872354676911,
    0.7415311855993945, -0.7415311855993945,
    0.8648644233597691, -0.8648644233597691,
    0.9491079123427585, -0.9491079123427585,
    0.9914553711208126, -0.9914553711208126,
};
__constant__ double kGkWeights[15] = {
    0.2094821410847278,
    0.2044329400752989, 0.2044329400752989,
    0.1903505780647854, 0.1903505780647854,
    0.1690047266392679, 0.1690047266392679,
    0.1406532597155259, 0.1406532597155259,
    0.1047900103222502, 0.1047900103222502,
    0.0630920926299786, 0.0630920926299786,
    0.0229353220105292, 0.0229353220105292,
};
__constant__ int kG7Idx[7] = {0, 3, 4, 7, 8, 11, 12};
__constant__ double kG7Weights[7] = {
    0.4179591836734694,
    0.3818300505051189, 0.3818300505051189,
    0.2797053914892766, 0.2797053914892766,
    0.1294849661688697, 0.1294849661688697,
};

// heston.rs:180-187: K15 and |K15 - G7| of one panel from its 15 node values,
// summed in index order.
__device__ void panel_sums(const double* f, double a, double b, double& kout, double& eout) {
    const double h = 0.5 * (b - a);
    double k = 0.0, g = 0.0;
    for (int i = 0; i < 15; ++i) k += kGkWeights[i] * f[i];
    for (int j = 0; j < 7; ++j) g += kG7Weights[j] * f[kG7Idx[j]];
    kout = k * h;
    eout = fabs(k - g) * h;
}

struct Panels {  // one adaptive integral's live panels, in the reference's order
    double e[kCfMaxPanels + 2];
    double k[kCfMaxPanels + 2];
    double a[kCfMaxPanels + 2];
    double b[kCfMaxPanels + 2];
};

// The derivative integrals' K15 sums per live panel (DERIV).
struct PanelsD {
    double k[3][kCfMaxPanels + 2];
};

// DERIV: Bates options also integrate the three jump-parameter derivatives of
// the integrands on the same panels, the value's error driving the
// bisection (ad.rs:639-670), and dout receives d(price)/d(lambda, mu_j,
// sigma_j) per option (ad.rs:711-730).  Options without the deriv flag
// (and Heston options) in a DERIV launch get zeros.
template <bool DERIV>
__global__ __launch_bounds__(64) void k_cf(const CfDesc* __restrict__ opts, double* __restrict__ out,
                                           CfStats* __restrict__ stats, double* __restrict__ dout) {
    __shared__ Panels P[2];
    __shared__ double fv[2][2][15];
    __shared__ double s_total[2], s_err[2], s_kw[2], s_ew[2], s_aw[2], s_m[2], s_bw[2];
    __shared__ double s_kc[2][2], s_ec[2][2];
    __shared__ int s_n[2];
    // derivative state (sized 1 without DERIV)
    constexpr int ND = DERIV ? 2 : 1;
    __shared__ PanelsD PD[ND];
    __shared__ double fd[ND][2][3][15];
    __shared__ double s_totd[ND][3], s_kdw[ND][3], s_kdc[ND][2][3];

    const CfDesc p = opts[blockIdx.x];
    const int lane = threadIdx.x;
    const int half = lane >> 5;  // 0: P1 (stock measure), 1: P2
    const int hl = lane & 31;
    const bool is_p1 = half == 0;
    const double x = log(p.spot / p.strike);
    // CF(-i), the P1 normaliser (heston.rs:91, bates.rs:55)
    const cpx norm = is_p1 ? model_cf(C(0.0, -1.0), p) : C(1.0, 0.0);
    Panels& Q = P[half];
    const bool deriv = DERIV && p.bates != 0 && p.deriv != 0;  // uniform per block
    const int hd = DERIV ? half : 0;

    // Round 0 evaluates the first panel, [0, 1], as a lone left child; every
    // later round splits the panel with the largest error into two children.
    // One evaluation site keeps a single inlined copy of the CF code.
    if (hl == 0) {
        s_aw[half] = 0.0;
        s_m[half] = 1.0;
        s_bw[half] = 1.0;
    }
    for (int round = 0;; ++round) {
        bool act = true;
        if (round > 0) {
            const bool act0 = s_err[0] > kCfTol && s_n[0] < kCfMaxPanels;
            const bool act1 = s_err[1] > kCfTol && s_n[1] < kCfMaxPanels;
            if (!act0 && !act1) break;
            act = half == 0 ? act0 : act1;
            if (act) {
                // the panel with the largest error, the last one on ties
                const int np = s_n[half];
                double bv = -1.0;
                int bi = -1;
                for (int j = hl; j < np; j += 32) {
                    const double v = Q.e[j];
                    if (v >= bv) {
                        bv = v;
                        bi = j;
                    }
                }
#pragma unroll
                for (int off = 16; off >= 1; off >>= 1) {
                    const double ov = __shfl_xor(bv, off, 32);
                    const int oi = __shfl_xor(bi, off, 32);
                    if (ov > bv || (ov == bv && oi > bi)) {
                        bv = ov;
                        bi = oi;
                    }
                }
                if (hl == 0) {
                    const int w = bi, last = np - 1;
                    s_ew[half] = Q.e[w];
                    s_kw[half] = Q.k[w];
                    if (deriv)
                        for (int k = 0; k < 3; ++k) {
                            s_kdw[hd][k] = PD[hd].k[k][w];
                            PD[hd].k[k][w] = PD[hd].k[k][last];
                        }
                    const double aw = Q.a[w], bw = Q.b[w];
                    Q.e[w] = Q.e[last];  // swap_remove
                    Q.k[w] = Q.k[last];
                    Q.a[w] = Q.a[last];
                    Q.b[w] = Q.b[last];
                    s_n[half] = last;
                    s_aw[half] = aw;
                    s_bw[half] = bw;
                    s_m[half] = 0.5 * (aw + bw);
                }
            }
        }
        __syncthreads();
        const int children = round == 0 ? 1 : 2;
        if (act && hl < 15 * children) {
            const int c = hl >= 15 ? 1 : 0, i = hl - 15 * c;
            const double a = c ? s_m[half] : s_aw[half];
            const double b = c ? s_bw[half] : s_m[half];
            const double cc = 0.5 * (a + b), h = 0.5 * (b - a);
            if (deriv) {
                double dv[3];
                fv[half][c][i] = mapped<true>(cc + h * kGkNodes[i], p, x, is_p1, norm, dv);
                for (int k = 0; k < 3; ++k) fd[hd][c][k][i] = dv[k];
            } else {
                fv[half][c][i] = mapped<false>(cc + h * kGkNodes[i], p, x, is_p1, norm, nullptr);
            }
        }
        __syncthreads();
        if (act && hl < children) {
            const double a = hl ? s_m[half] : s_aw[half];
            const double b = hl ? s_bw[half] : s_m[half];
            panel_sums(fv[half][hl], a, b, s_kc[half][hl], s_ec[half][hl]);
            if (deriv) {  // ad.rs:639-647: K15 of the derivative values, times h
                const double h = 0.5 * (b - a);
                for (int k = 0; k < 3; ++k) {
                    double kd = 0.0;
                    for (int i = 0; i < 15; ++i) kd += kGkWeights[i] * fd[hd][hl][k][i];
                    s_kdc[hd][hl][k] = kd * h;
                }
            }
        }
        __syncthreads();
        if (act && hl == 0) {
            const double kl = s_kc[half][0], el = s_ec[half][0];
            if (deriv)
                for (int k = 0; k < 3; ++k) {
                    const double dl = s_kdc[hd][0][k];
                    if (round == 0) {
                        PD[hd].k[k][0] = dl;
                        s_totd[hd][k] = dl;
                    } else {
                        const double dr = s_kdc[hd][1][k];
                        s_totd[hd][k] += dl + dr - s_kdw[hd][k];
                        PD[hd].k[k][s_n[half]] = dl;
                        PD[hd].k[k][s_n[half] + 1] = dr;
                    }
                }
            if (round == 0) {
                Q.e[0] = el;
                Q.k[0] = kl;
                Q.a[0] = 0.0;
                Q.b[0] = 1.0;
                s_n[half] = 1;
                s_total[half] = kl;
                s_err[half] = el;
            } else {
                const double kr = s_kc[half][1], er = s_ec[half][1];
                s_total[half] += kl + kr - s_kw[half];
                s_err[half] += el + er - s_ew[half];
                int n = s_n[half];
                Q.e[n] = el;
                Q.k[n] = kl;
                Q.a[n] = s_aw[half];
                Q.b[n] = s_m[half];
                ++n;
                Q.e[n] = er;
                Q.k[n] = kr;
                Q.a[n] = s_m[half];
                Q.b[n] = s_bw[half];
                s_n[half] = n + 1;
            }
        }
        __syncthreads();
    }

    if (lane == 0) {
        // heston.rs:94-98, put by parity (heston.rs:68-71)
        const double PI = 3.141592653589793;
        const double p1 = 0.5 + s_total[0] / PI;
        const double p2 = 0.5 + s_total[1] / PI;
        const double eq = exp(-p.div_yield * p.expiry), er = exp(-p.rate * p.expiry);
        double v = p.spot * eq * p1 - p.strike * er * p2;
        v = v > 0.0 ? v : 0.0;
        if (p.put) v = v - p.spot * eq + p.strike * er;
        out[blockIdx.x] = v;
        if (stats) stats[blockIdx.x] = CfStats{(uint32_t)s_n[0], (uint32_t)s_n[1]};
        if (DERIV) {
            // ad.rs:711-730: seq dP1 - ker dP2, the same for a put (parity terms
            // do not depend on the jump parameters)
            for (int k = 0; k < 3; ++k) {
                double dk = 0.0;
                if (deriv) dk = p.spot * eq * (s_totd[0][k] / PI) - p.strike * er * (s_totd[1][k] / PI);
                dout[3 * blockIdx.x + k] = dk;
            }
        }
    }
}

}  // namespace

hipError_t launch_cf(const CfDesc* opts, uint32_t n, double* out, CfStats* stats, double* dout, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (dout)
        hipLaunchKernelGGL(k_cf<true>, dim3(n), dim3(64), 0, s, opts, out, stats, dout);
    else
        hipLaunchKernelGGL(k_cf<false>, dim3(n), dim3(64), 0, s, opts, out, stats, dout);
    return hipGetLastError();
}

}  // namespace ope
