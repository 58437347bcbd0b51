/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Statistical oracle for the Monte Carlo path: C restatement of the
 * reference's analytic pricers, which its own MC tests compare against
 * (mc.rs:393-429):
 *   heston_price   heston.rs:61-72, heston_call 82-99, cf_integrand 101-109,
 *                  stable_cf (Albrecher form) 113-127, fast_csqrt 155-175,
 *                  gk15_panel 180-187, adaptive_gk 193-212, gk_integrate 220-229
 *   bates_price    bates.rs:29-39, bates_call 49-63, bates_integrand 65-73, jump_cf 77-83
 *   bates_jump_sens  ad.rs:882-895 (forward_pass 695-731 with the jump
 *                  parameter active, jump_factor 218-234, dual_integrand
 *                  609-630, adaptive_gk_dual 639-684): the price and its
 *                  derivatives in lambda, mu_j, sigma_j, each derivative
 *                  integrand (the exact derivative of the jump factor times
 *                  the integrand) summed on the panels the value's error
 *                  chooses
 *   bsm_price      bsm.rs:7-12, 55-60
 * plus the textbook Merton (1976) jump-diffusion series, which has no
 * reference counterpart and pins the Bates MC at zero vol-of-vol.
 *
 * Complex arithmetic follows num-complex 0.4 (naive product and quotient,
 * exp via from_polar, ln via (ln|z|, atan2)).  Compiled -ffp-contract=off.
 */
#include "oracle.h"

#include <math.h>

typedef struct { double re, im; } cpx;

double ora_ncdf(double x);
double ora_bsm_price(double s, double k, double t, double r, double q, double vol, int opt_type);

static inline cpx C(double re, double im) { cpx z = {re, im}; return z; }
static inline cpx cadd(cpx a, cpx b) { return C(a.re + b.re, a.im + b.im); }
static inline cpx csub(cpx a, cpx b) { return C(a.re - b.re, a.im - b.im); }
static inline cpx cmul(cpx a, cpx b) { return C(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline cpx cdiv(cpx a, cpx b) {
    double n = b.re * b.re + b.im * b.im;
    double re = a.re * b.re + a.im * b.im;
    double im = a.im * b.re - a.re * b.im;
    return C(re / n, im / n);
}
static inline cpx cscale(cpx a, double s) { return C(a.re * s, a.im * s); }       /* Complex * f64 */
static inline cpx cradd(cpx a, double s) { return C(a.re + s, a.im); }            /* Complex + f64 */
static inline cpx crsub_l(double s, cpx a) { return C(s - a.re, -a.im); }         /* f64 - Complex */
static inline cpx cexp_(cpx z) { double e = exp(z.re); return C(e * cos(z.im), e * sin(z.im)); }
static inline cpx clog_(cpx z) { return C(log(hypot(z.re, z.im)), atan2(z.im, z.re)); }

/* heston.rs:155-175 */
static cpx fast_csqrt(cpx z) {
    double a = z.re, b = z.im;
    if (b == 0.0) {
        if (a >= 0.0) return C(sqrt(a), b);
        double m = sqrt(-a);
        return C(0.0, signbit(b) ? -m : m);
    }
    double r = hypot(a, b);
    if (a >= 0.0) {
        double re_out = sqrt((r + a) / 2.0);
        double im_out = b / (2.0 * re_out);
        return C(re_out, im_out);
    } else {
        double im_out = copysign(sqrt((r - a) / 2.0), b);
        double re_out = b / (2.0 * im_out);
        return C(re_out, im_out);
    }
}

/* heston.rs:113-127 */
static cpx stable_cf(cpx phi, double t, double r, const ora_heston *p) {
    const cpx I = C(0.0, 1.0);
    double v0 = p->v0, kappa = p->kappa, theta = p->theta, sigma = p->sigma, rho = p->rho;
    cpx xi = crsub_l(kappa, cmul(cscale(phi, rho * sigma), I));
    cpx d = fast_csqrt(cadd(cmul(xi, xi), cmul(cscale(phi, sigma * sigma), cadd(phi, I))));
    cpx g = cdiv(csub(xi, d), cadd(xi, d));
    cpx edt = cexp_(cscale(C(-d.re, -d.im), t));
    cpx a = cdiv(cradd(cmul(g, edt), -1.0), cradd(g, -1.0));
    cpx c = cscale(csub(cscale(csub(xi, d), t), cscale(clog_(a), 2.0)), kappa * theta / (sigma * sigma));
    cpx dd = cdiv(cmul(cscale(csub(xi, d), v0), crsub_l(1.0, edt)),
                  cscale(crsub_l(1.0, cmul(g, edt)), sigma * sigma));
    cpx rpi = cscale(cmul(cscale(phi, r), I), t);
    return cexp_(cadd(cadd(rpi, c), dd));
}

/* bates.rs:77-83 */
static cpx jump_cf(cpx phi, double t, double lambda, double mu_j, double sigma_j) {
    const cpx I = C(0.0, 1.0);
    double comp = exp(mu_j + 0.5 * sigma_j * sigma_j) - 1.0;
    cpx jump = cexp_(csub(cscale(cmul(phi, I), mu_j),
                          cscale(cscale(cmul(cscale(phi, 0.5), phi), sigma_j), sigma_j)));
    cpx inner = csub(cradd(jump, -1.0), cscale(cmul(I, phi), comp));
    return cexp_(cscale(inner, lambda * t));
}

typedef struct {
    double x, t, r;
    const ora_heston *p;
    int is_p1;
    int has_norm;
    cpx norm;
    int bates;
    double lambda, mu_j, sigma_j;
} integrand_ctx;

/* heston.rs:101-109 and bates.rs:65-73 */
static double integrand(const integrand_ctx *c, double u) {
    cpx phi = c->is_p1 ? C(u, -1.0) : C(u, 0.0);
    cpx cf = stable_cf(phi, c->t, c->r, c->p);
    if (c->bates) cf = cmul(cf, jump_cf(phi, c->t, c->lambda, c->mu_j, c->sigma_j));
    if (c->has_norm) cf = cdiv(cf, c->norm);
    cpx num = cmul(cexp_(C(0.0, u * c->x)), cf);
    return cdiv(num, C(0.0, u)).re;
}

/* heston.rs:28-59: GK15 nodes and weights, G7 subset. */
static const double GK_NODES[15] = {
    0.0,
    0.2077849550078985, -0.2077849550078985,
    0.4058451513773972, -0.4058451513773972,
    0.5860872354676911, -0.5860872354676911,
    0.7415311855993945, -0.7415311855993945,
    0.8648644233597691, -0.8648644233597691,
    0.9491079123427585, -0.9491079123427585,
    0.9914553711208126, -0.9914553711208126,
};
static const double GK_WEIGHTS[15] = {
    0.2094821410847278,
    0.2044329400752989, 0.2044329400752989,
    0.1903505780647854, 0.1903505780647854,
    0.1690047266392679, 0.1690047266392679,
    0.1406532597155259, 0.1406532597155259,
    0.1047900103222502, 0.1047900103222502,
    0.0630920926299786, 0.0630920926299786,
    0.0229353220105292, 0.0229353220105292,
};
static const int G7_IDX[7] = {0, 3, 4, 7, 8, 11, 12};
static const double G7_WEIGHTS[7] = {
    0.4179591836734694,
    0.3818300505051189, 0.3818300505051189,
    0.2797053914892766, 0.2797053914892766,
    0.1294849661688697, 0.1294849661688697,
};

/* heston.rs:220-229: the u = (1-t)/t map of [0, inf) onto [0, 1]. */
static double mapped(const integrand_ctx *c, double t) {
    if (t <= 0.0) return 0.0;
    double u = (1.0 - t) / t;
    if (u < 1e-12) return 0.0;
    double v = integrand(c, u) / (t * t);
    return isfinite(v) ? v : 0.0;
}

/* heston.rs:180-187 */
static void gk15_panel(const integrand_ctx *c, double a, double b, double *kout, double *eout) {
    double cc = 0.5 * (a + b), h = 0.5 * (b - a);
    double fv[15];
    for (int i = 0; i < 15; ++i) fv[i] = mapped(c, cc + h * GK_NODES[i]);
    double k = 0.0, g = 0.0;
    for (int i = 0; i < 15; ++i) k += GK_WEIGHTS[i] * fv[i];
    for (int j = 0; j < 7; ++j) g += G7_WEIGHTS[j] * fv[G7_IDX[j]];
    *kout = k * h;
    *eout = fabs(k - g) * h;
}

/* heston.rs:193-212: bisect the panel with the largest error (the last such
 * panel on ties, as Iterator::max_by does) until the error sum <= tol or 200
 * panels; swap_remove semantics kept. */
static double adaptive_gk(const integrand_ctx *c, double a, double b, double tol) {
    enum { MAX_PANELS = 200 };
    double pe[MAX_PANELS + 2], pk[MAX_PANELS + 2], pa[MAX_PANELS + 2], pb[MAX_PANELS + 2];
    int n = 0;
    double k0, e0;
    gk15_panel(c, a, b, &k0, &e0);
    pe[0] = e0; pk[0] = k0; pa[0] = a; pb[0] = b; n = 1;
    double total = k0, err = e0;
    while (err > tol && n < MAX_PANELS) {
        int w = 0;
        for (int i = 1; i < n; ++i) if (pe[i] >= pe[w]) w = i;
        double ew = pe[w], kw = pk[w], aw = pa[w], bw = pb[w];
        /* swap_remove */
        pe[w] = pe[n - 1]; pk[w] = pk[n - 1]; pa[w] = pa[n - 1]; pb[w] = pb[n - 1];
        n -= 1;
        double m = 0.5 * (aw + bw);
        double kl, el, kr, er;
        gk15_panel(c, aw, m, &kl, &el);
        gk15_panel(c, m, bw, &kr, &er);
        total += kl + kr - kw;
        err += el + er - ew;
        pe[n] = el; pk[n] = kl; pa[n] = aw; pb[n] = m; n++;
        pe[n] = er; pk[n] = kr; pa[n] = m; pb[n] = bw; n++;
    }
    return total;
}

static double cf_call(double s, double k, double t, double r, double q, const ora_heston *p, int bates,
                      double lambda, double mu_j, double sigma_j) {
    integrand_ctx c;
    c.x = log(s / k);
    c.t = t;
    c.r = r;
    c.p = p;
    c.bates = bates;
    c.lambda = lambda;
    c.mu_j = mu_j;
    c.sigma_j = sigma_j;
    cpx mi = C(0.0, -1.0);
    cpx cf_mi = stable_cf(mi, t, r, p);
    if (bates) cf_mi = cmul(cf_mi, jump_cf(mi, t, lambda, mu_j, sigma_j));
    c.is_p1 = 1; c.has_norm = 1; c.norm = cf_mi;
    double i1 = adaptive_gk(&c, 0.0, 1.0, 1e-8);
    c.is_p1 = 0; c.has_norm = 0;
    double i2 = adaptive_gk(&c, 0.0, 1.0, 1e-8);
    const double PI = 3.141592653589793;
    double p1 = 0.5 + i1 / PI;
    double p2 = 0.5 + i2 / PI;
    double v = s * exp(-q * t) * p1 - k * exp(-r * t) * p2;
    return v > 0.0 ? v : 0.0;
}

/* heston.rs:61-72 (put by parity) */
double ora_heston_price(double s, double k, double t, double r, double q, const ora_heston *p, int opt_type) {
    double call = cf_call(s, k, t, r, q, p, 0, 0, 0, 0);
    if (opt_type == ORA_CALL) return call;
    return call - s * exp(-q * t) + k * exp(-r * t);
}

/* bates.rs:29-39 */
double ora_bates_price(double s, double k, double t, double r, double q, const ora_heston *p, double lambda,
                       double mu_j, double sigma_j, int opt_type) {
    double call = cf_call(s, k, t, r, q, p, 1, lambda, mu_j, sigma_j);
    if (opt_type == ORA_CALL) return call;
    return call - s * exp(-q * t) + k * exp(-r * t);
}

/* ---------------------------------------------- jump-parameter derivatives */

/* d ln J(phi) / d(lambda, mu_j, sigma_j) of jump_cf: lambda T (E - 1 - i phi k)
 * with E = exp(i phi mu_j - sigma_j^2 phi^2 / 2), k = e^(mu_j + sigma_j^2/2) - 1,
 * differentiated by hand (ad.rs:218-234 propagates the same with dual numbers). */
static void jump_dlog(cpx phi, double t, double lambda, double mu_j, double sigma_j, cpx g[3]) {
    const cpx I = C(0.0, 1.0);
    double ec = exp(mu_j + 0.5 * sigma_j * sigma_j);
    cpx E = cexp_(csub(cscale(cmul(phi, I), mu_j), cscale(cscale(cmul(cscale(phi, 0.5), phi), sigma_j), sigma_j)));
    cpx iphi = cmul(I, phi);
    cpx inner = csub(cradd(E, -1.0), cscale(iphi, ec - 1.0));
    g[0] = cscale(inner, t);
    g[1] = cscale(cmul(iphi, cradd(E, -ec)), lambda * t);
    g[2] = cscale(cadd(cmul(cmul(phi, phi), E), cscale(iphi, ec)), -lambda * t * sigma_j);
}

/* integrand and its three jump derivatives on the t map; the value keeps the
 * pricer's guard, a derivative is 0 where it or the value is not finite */
static double mapped_d(const integrand_ctx *c, double t, double d[3]) {
    d[0] = d[1] = d[2] = 0.0;
    if (t <= 0.0) return 0.0;
    double u = (1.0 - t) / t;
    if (u < 1e-12) return 0.0;
    cpx phi = c->is_p1 ? C(u, -1.0) : C(u, 0.0);
    cpx cf = cmul(stable_cf(phi, c->t, c->r, c->p), jump_cf(phi, c->t, c->lambda, c->mu_j, c->sigma_j));
    if (c->has_norm) cf = cdiv(cf, c->norm);
    cpx z = cdiv(cmul(cexp_(C(0.0, u * c->x)), cf), C(0.0, u));
    double v = z.re / (t * t);
    if (!isfinite(v)) return 0.0;
    cpx g[3];
    jump_dlog(phi, c->t, c->lambda, c->mu_j, c->sigma_j, g);
    for (int k = 0; k < 3; ++k) {
        double dk = cmul(z, g[k]).re / (t * t);
        d[k] = isfinite(dk) ? dk : 0.0;
    }
    return v;
}

static void gk15_panel_d(const integrand_ctx *c, double a, double b, double *kout, double *eout, double kd[3]) {
    double cc = 0.5 * (a + b), h = 0.5 * (b - a);
    double fv[15], fd[15][3];
    for (int i = 0; i < 15; ++i) fv[i] = mapped_d(c, cc + h * GK_NODES[i], fd[i]);
    double k = 0.0, g = 0.0;
    for (int i = 0; i < 15; ++i) k += GK_WEIGHTS[i] * fv[i];
    for (int j = 0; j < 7; ++j) g += G7_WEIGHTS[j] * fv[G7_IDX[j]];
    for (int q = 0; q < 3; ++q) {
        double s = 0.0;
        for (int i = 0; i < 15; ++i) s += GK_WEIGHTS[i] * fd[i][q];
        kd[q] = s * h;
    }
    *kout = k * h;
    *eout = fabs(k - g) * h;
}

/* adaptive_gk carrying the three derivative sums on the same panels */
static double adaptive_gk_d(const integrand_ctx *c, double tol, double dtot[3]) {
    enum { MAX_PANELS = 200 };
    double pe[MAX_PANELS + 2], pk[MAX_PANELS + 2], pa[MAX_PANELS + 2], pb[MAX_PANELS + 2];
    double pd[MAX_PANELS + 2][3];
    int n = 0;
    double k0, e0, d0[3];
    gk15_panel_d(c, 0.0, 1.0, &k0, &e0, d0);
    pe[0] = e0; pk[0] = k0; pa[0] = 0.0; pb[0] = 1.0; n = 1;
    for (int q = 0; q < 3; ++q) { pd[0][q] = d0[q]; dtot[q] = d0[q]; }
    double total = k0, err = e0;
    while (err > tol && n < MAX_PANELS) {
        int w = 0;
        for (int i = 1; i < n; ++i) if (pe[i] >= pe[w]) w = i;
        double ew = pe[w], kw = pk[w], aw = pa[w], bw = pb[w], dw[3];
        for (int q = 0; q < 3; ++q) { dw[q] = pd[w][q]; pd[w][q] = pd[n - 1][q]; }
        pe[w] = pe[n - 1]; pk[w] = pk[n - 1]; pa[w] = pa[n - 1]; pb[w] = pb[n - 1];
        n -= 1;
        double m = 0.5 * (aw + bw);
        double kl, el, kr, er, dl[3], dr[3];
        gk15_panel_d(c, aw, m, &kl, &el, dl);
        gk15_panel_d(c, m, bw, &kr, &er, dr);
        total += kl + kr - kw;
        err += el + er - ew;
        for (int q = 0; q < 3; ++q) dtot[q] += dl[q] + dr[q] - dw[q];
        pe[n] = el; pk[n] = kl; pa[n] = aw; pb[n] = m;
        for (int q = 0; q < 3; ++q) pd[n][q] = dl[q];
        n++;
        pe[n] = er; pk[n] = kr; pa[n] = m; pb[n] = bw;
        for (int q = 0; q < 3; ++q) pd[n][q] = dr[q];
        n++;
    }
    return total;
}

/* ad.rs:882-895: out = {price, d/d lambda, d/d mu_j, d/d sigma_j} */
void ora_bates_jump_sens(double s, double k, double t, double r, double q, const ora_heston *p, double lambda,
                         double mu_j, double sigma_j, int opt_type, double *out) {
    integrand_ctx c;
    c.x = log(s / k);
    c.t = t;
    c.r = r;
    c.p = p;
    c.bates = 1;
    c.lambda = lambda;
    c.mu_j = mu_j;
    c.sigma_j = sigma_j;
    cpx mi = C(0.0, -1.0);
    c.is_p1 = 1; c.has_norm = 1; c.norm = cmul(stable_cf(mi, t, r, p), jump_cf(mi, t, lambda, mu_j, sigma_j));
    double d1[3], d2[3];
    double i1 = adaptive_gk_d(&c, 1e-8, d1);
    c.is_p1 = 0; c.has_norm = 0;
    double i2 = adaptive_gk_d(&c, 1e-8, d2);
    const double PI = 3.141592653589793;
    double seq = s * exp(-q * t), ker = k * exp(-r * t);
    double call = seq * (0.5 + i1 / PI) - ker * (0.5 + i2 / PI);
    call = call > 0.0 ? call : 0.0;
    out[0] = opt_type == ORA_CALL ? call : call - seq + ker;
    for (int j = 0; j < 3; ++j) out[1 + j] = seq * (d1[j] / PI) - ker * (d2[j] / PI);
}

/* n options, 15 doubles each: spot, strike, expiry, rate, div_yield, v0,
 * kappa, theta, sigma, rho, lambda, mu_j, sigma_j, model (0 Heston, 1 Bates),
 * opt_type; the loop a batch.rs chain runs per thread (batch.rs:23-42). */
void ora_cf_price_many(const double *args, int64_t n, double *out) {
    for (int64_t i = 0; i < n; ++i) {
        const double *a = args + 15 * i;
        ora_heston h = {a[5], a[6], a[7], a[8], a[9]};
        const int opt = (int)a[14];
        out[i] = a[13] != 0.0 ? ora_bates_price(a[0], a[1], a[2], a[3], a[4], &h, a[10], a[11], a[12], opt)
                              : ora_heston_price(a[0], a[1], a[2], a[3], a[4], &h, opt);
    }
}

/* greeks.rs:19-82: bump-and-reprice Greeks through the CF pricers, 14 prices
 * per option (dS = 1 % of spot, dv = 1 vol point on sqrt(v0), dr = 1 bp,
 * dt = 1 day).  out: price, delta, gamma, vega, theta, rho, vanna, volga. */
static double cf_price(double s, double k, double t, double r, double q, const ora_heston *p, int bates,
                       double lambda, double mu_j, double sigma_j, int opt) {
    return bates ? ora_bates_price(s, k, t, r, q, p, lambda, mu_j, sigma_j, opt) : ora_heston_price(s, k, t, r, q, p, opt);
}

static ora_heston with_v0(const ora_heston *p, double v0) {
    ora_heston h = *p;
    h.v0 = v0;
    return h;
}

void ora_cf_greeks(double s, double k, double t, double r, double q, const ora_heston *p, int bates, double lambda,
                   double mu_j, double sigma_j, int opt, double *out) {
#define PX(S, K, T, R, Q, P) cf_price(S, K, T, R, Q, P, bates, lambda, mu_j, sigma_j, opt)
    double price = PX(s, k, t, r, q, p);
    double ds = 0.01 * s, dv = 0.01, dr = 1e-4, dt = 1.0 / 365.0;
    double pu = PX(s + ds, k, t, r, q, p), pd = PX(s - ds, k, t, r, q, p);
    double delta = (pu - pd) / (2.0 * ds);
    double gamma = (pu - 2.0 * price + pd) / (ds * ds);
    double v_cur = sqrt(p->v0);
    double vu = v_cur + dv, vd = fmax(v_cur - dv, 1e-8);
    ora_heston p_vup = with_v0(p, vu * vu), p_vdn = with_v0(p, vd * vd);
    double vega = (PX(s, k, t, r, q, &p_vup) - PX(s, k, t, r, q, &p_vdn)) / (2.0 * dv);
    double t_dn = fmax(t - dt, 1e-6);
    double theta = (PX(s, k, t_dn, r, q, p) - price) / dt;
    double rho = (PX(s, k, t, r + dr, q, p) - PX(s, k, t, r - dr, q, p)) / (2.0 * dr);
    double delta_vup = (PX(s + ds, k, t, r, q, &p_vup) - PX(s - ds, k, t, r, q, &p_vup)) / (2.0 * ds);
    double delta_vdn = (PX(s + ds, k, t, r, q, &p_vdn) - PX(s - ds, k, t, r, q, &p_vdn)) / (2.0 * ds);
    double vanna = (delta_vup - delta_vdn) / (2.0 * dv);
    double vu2 = v_cur + 2.0 * dv, vd2 = fmax(v_cur - 2.0 * dv, 1e-8);
    ora_heston p_vup2 = with_v0(p, vu2 * vu2), p_vdn2 = with_v0(p, vd2 * vd2);
    double vega_up = (PX(s, k, t, r, q, &p_vup2) - price) / (2.0 * dv);
    double vega_dn = (price - PX(s, k, t, r, q, &p_vdn2)) / (2.0 * dv);
    double volga = (vega_up - vega_dn) / (2.0 * dv);
#undef PX
    out[0] = price; out[1] = delta; out[2] = gamma; out[3] = vega;
    out[4] = theta; out[5] = rho; out[6] = vanna; out[7] = volga;
}

double ora_ncdf(double x);

/* math.rs:8-13 */
static double npdf(double x) { return 0.3989422804014327 * exp(-0.5 * x * x); }

/* bsm.rs:11-45: price and closed-form Greeks.  out: price, delta, gamma,
 * vega, theta, rho, vanna, volga. */
void ora_bsm_greeks(double s, double k, double t, double r, double q, double v, int opt_type, double *out) {
    double vt = v * sqrt(t);
    double d1 = (log(s / k) + (r - q + 0.5 * v * v) * t) / vt;
    double d2 = d1 - vt;
    double phi = opt_type == ORA_CALL ? 1.0 : -1.0;
    double nd1 = ora_ncdf(phi * d1), nd2 = ora_ncdf(phi * d2), npd1 = npdf(d1);
    double eq = exp(-q * t), er = exp(-r * t);
    double seq = s * eq, ker = k * er, sqt = sqrt(t);
    out[0] = phi * (seq * nd1 - ker * nd2);
    out[1] = phi * eq * nd1;
    out[2] = eq * npd1 / (s * v * sqt);
    out[3] = seq * npd1 * sqt;
    out[4] = -seq * npd1 * (v * sqt) / (2.0 * t) - phi * (q * seq * nd1 - r * ker * nd2);
    out[5] = phi * k * t * er * nd2;
    out[6] = -eq * npd1 * d2 / v;
    out[7] = out[3] * d1 * d2 / v;
}

/* iv.rs:14-117: bounds check, Brenner-Subrahmanyam guess (or bisection),
 * Halley.  Returns 1 and *vol when a volatility is found, else 0. */
static double bsm_at(double s, double k, double t, double r, double q, double v, int opt) {
    return ora_bsm_price(s, k, t, r, q, v, opt);
}

static int iv_bisect(double s, double k, double t, double r, double q, int opt, double target, double *vol) {
    double lo = 1e-4, hi = 10.0;
    double flo = bsm_at(s, k, t, r, q, lo, opt) - target, fhi = bsm_at(s, k, t, r, q, hi, opt) - target;
    if (flo * fhi > 0.0) return 0;
    for (int i = 0; i < 60; ++i) {
        double mid = 0.5 * (lo + hi);
        if (bsm_at(s, k, t, r, q, mid, opt) - target < 0.0) lo = mid; else hi = mid;
        if (hi - lo < 1e-9) { *vol = mid; return 1; }
    }
    *vol = 0.5 * (lo + hi);
    return 1;
}

int ora_implied_vol(double s, double k, double t, double r, double q, int opt, double price, double *vol) {
    /* price_in_bounds (iv.rs:22-32) */
    if (price <= 0.0) return 0;
    double er = exp(-r * t), eq = exp(-q * t);
    double intrinsic = opt == ORA_CALL ? fmax(s * eq - k * er, 0.0) : fmax(k * er - s * eq, 0.0);
    if (!(price >= intrinsic)) return 0;
    /* initial_guess (iv.rs:37-50) */
    double fwd = s * eq / er;
    double bs = price * sqrt(2.0 * 3.141592653589793 / t) / (fwd * er);
    double x = log(fwd / k);
    double v = bs / fmax(exp(-0.5 * x * x), 0.01);
    v = v < 0.001 ? 0.001 : (v > 10.0 ? 10.0 : v);
    if (!(fabs(bsm_at(s, k, t, r, q, v, opt) - price) < 0.2 * price)) {
        if (!iv_bisect(s, k, t, r, q, opt, price, &v)) return 0;
    }
    /* halley_solve (iv.rs:68-90) with vega_and_volga (93-103) */
    for (int it = 0; it < 10; ++it) {
        double err = bsm_at(s, k, t, r, q, v, opt) - price;
        if (fabs(err) < 1e-10) { *vol = v; return 1; }
        double vt = v * sqrt(t);
        double d1 = (log(s / k) + (r - q + 0.5 * v * v) * t) / vt;
        double d2 = d1 - vt;
        double vega = s * eq * npdf(d1) * sqrt(t);
        double volga = vega * d1 * d2 / v;
        if (fabs(vega) < 1e-14) return 0;
        double denom = 1.0 - err * volga / (2.0 * vega * vega);
        denom = denom < 0.5 ? 0.5 : (denom > 2.0 ? 2.0 : denom);
        v -= err / (vega * denom);
        if (v <= 0.0) v = 1e-8;
    }
    if (fabs(bsm_at(s, k, t, r, q, v, opt) - price) < 1e-6) { *vol = v; return 1; }
    return 0;
}

/* Chains of the two above, the loops batch.rs:13-15 and 51-56 run per
 * thread.  args rows: spot, strike, expiry, rate, div_yield, vol, opt_type
 * (Greeks) or spot, strike, expiry, rate, div_yield, opt_type, price (IV);
 * an implied volatility that is not found is NaN. */
void ora_bsm_greeks_many(const double *args, int64_t n, double *out) {
    for (int64_t i = 0; i < n; ++i) {
        const double *a = args + 7 * i;
        ora_bsm_greeks(a[0], a[1], a[2], a[3], a[4], a[5], (int)a[6], out + 8 * i);
    }
}

void ora_implied_vol_many(const double *args, int64_t n, double *vol) {
    for (int64_t i = 0; i < n; ++i) {
        const double *a = args + 7 * i;
        double v;
        vol[i] = ora_implied_vol(a[0], a[1], a[2], a[3], a[4], (int)a[5], a[6], &v) ? v : NAN;
    }
}

/* bsm.rs:7-12, 55-60 */
double ora_bsm_price(double s, double k, double t, double r, double q, double vol, int opt_type) {
    double vt = vol * sqrt(t);
    double d1 = (log(s / k) + (r - q + 0.5 * vol * vol) * t) / vt;
    double d2 = d1 - vt;
    double phi = opt_type == ORA_CALL ? 1.0 : -1.0;
    return phi * (s * exp(-q * t) * ora_ncdf(phi * d1) - k * exp(-r * t) * ora_ncdf(phi * d2));
}

/* Merton (1976) lognormal jump-diffusion, constant diffusion variance `var`:
 * conditional on n jumps ln S_T is normal with forward
 * F_n = S exp((r_n - q) T), r_n = r - lambda kbar + n (mu_j + sigma_j^2/2) / T,
 * and total variance var T + n sigma_j^2, so the call is
 * sum_n Poisson(n; lambda T) e^{-rT} (F_n N(d1) - K N(d2)). */
double ora_merton_price(double s, double k, double t, double r, double q, double var, double lambda,
                        double mu_j, double sigma_j, int opt_type) {
    double kbar = exp(mu_j + 0.5 * sigma_j * sigma_j) - 1.0;
    double lp = lambda * t;
    double total = 0.0, w = exp(-lp);
    for (int n = 0; n < 200; ++n) {
        if (n > 0) w *= lp / n;
        double rn = r - lambda * kbar + n * (mu_j + 0.5 * sigma_j * sigma_j) / t;
        double vn = sqrt(var + n * sigma_j * sigma_j / t);
        /* BSM with rate rn for the forward, discounted at r */
        double fwd = s * exp((rn - q) * t);
        double sd = vn * sqrt(t);
        double d1 = (log(fwd / k) + 0.5 * sd * sd) / sd;
        double d2 = d1 - sd;
        double phi = opt_type == ORA_CALL ? 1.0 : -1.0;
        double price = exp(-r * t) * phi * (fwd * ora_ncdf(phi * d1) - k * ora_ncdf(phi * d2));
        total += w * price;
        if (n > 10 && w < 1e-18) break;
    }
    return total;
}
