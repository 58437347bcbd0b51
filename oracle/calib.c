/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * C restatement of the reference's calibration evaluations
 * (calibration.rs): the per-quote implied-volatility residual (573-587), the
 * central finite-difference Jacobian with bump points clamped to the
 * parameter bounds (613-652), and J'J / J'r (655-667).  Quotes are rows of
 * 8 doubles: spot, strike, expiry, rate, div_yield, opt_type, iv_market,
 * weight.  model 0: Heston (5 parameters, calibration.rs:104-141), 1: Bates
 * (8, 143-197).  Bates' three jump columns come from forward AD when every
 * quote's solver succeeds and its vega is at least 1e-10 (calibration.rs:
 * 186-199, 601-634): d(price)/d(param) from ora_bates_jump_sens (ad.rs:882)
 * divided by the BSM vega at the model's implied vol; otherwise the column is
 * differenced like the others.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>

double ora_heston_price(double s, double k, double t, double r, double q, const ora_heston *p, int opt_type);
double ora_bates_price(double s, double k, double t, double r, double q, const ora_heston *p, double lambda,
                       double mu_j, double sigma_j, int opt_type);
int ora_implied_vol(double s, double k, double t, double r, double q, int opt, double price, double *vol);
void ora_bates_jump_sens(double s, double k, double t, double r, double q, const ora_heston *p, double lambda,
                         double mu_j, double sigma_j, int opt_type, double *out);
void ora_bsm_greeks(double s, double k, double t, double r, double q, double vol, int opt_type, double *out);

static const double BUMP[8] = {1e-4, 1e-3, 1e-4, 1e-3, 1e-4, 1e-4, 1e-4, 1e-4};
static const double LO[8] = {1e-8, 1e-6, 1e-8, 1e-6, -0.9999, 0.0, -2.0, 1e-6};
static const double HI[8] = {5.0, 50.0, 5.0, 10.0, 0.9999, 10.0, 2.0, 2.0};

static double model_price(const double *q, int model, const double *p) {
    ora_heston h = {p[0], p[1], p[2], p[3], p[4]};
    return model ? ora_bates_price(q[0], q[1], q[2], q[3], q[4], &h, p[5], p[6], p[7], (int)q[5])
                 : ora_heston_price(q[0], q[1], q[2], q[3], q[4], &h, (int)q[5]);
}

/* calibration.rs:573-581 */
static double single_residual(const double *q, int model, const double *p) {
    double px = model_price(q, model, p), iv;
    return ora_implied_vol(q[0], q[1], q[2], q[3], q[4], (int)q[5], px, &iv) ? iv - q[6] : 0.0;
}

void ora_calib_residuals(const double *quotes, int64_t n, int model, const double *params, double *res) {
    for (int64_t i = 0; i < n; ++i) res[i] = single_residual(quotes + 8 * i, model, params);
}

/* calibration.rs:601-609 model_iv_and_vega, then the AD column entry
 * d(price)/d(param) / vega (calibration.rs:621-626); 0 when the solver finds
 * no volatility or |vega| < 1e-10 (the caller then differences the column). */
static int ad_entry(const double *q, const double *p, int col, double *out) {
    ora_heston h = {p[0], p[1], p[2], p[3], p[4]};
    double sens[4], iv, g[8];
    ora_bates_jump_sens(q[0], q[1], q[2], q[3], q[4], &h, p[5], p[6], p[7], (int)q[5], sens);
    if (!ora_implied_vol(q[0], q[1], q[2], q[3], q[4], (int)q[5], sens[0], &iv)) return 0;
    ora_bsm_greeks(q[0], q[1], q[2], q[3], q[4], iv, (int)q[5], g);
    if (fabs(g[3]) < 1e-10) return 0;
    *out = sens[1 + (col - 5)] / g[3];
    return 1;
}

void ora_calib_jacobian(const double *quotes, int64_t n, int model, const double *params, double *jac) {
    const int dim = model ? 8 : 5;
    double vu[8], vd[8];
    for (int64_t i = 0; i < n * dim; ++i) jac[i] = 0.0;
    for (int col = 0; col < dim; ++col) {
        if (model && col >= 5) {  /* AD for the whole column, or differences for all of it */
            int ok = 1;
            for (int64_t i = 0; ok && i < n; ++i) ok = ad_entry(quotes + 8 * i, params, col, &jac[i * dim + col]);
            if (ok) continue;
            for (int64_t i = 0; i < n; ++i) jac[i * dim + col] = 0.0;
        }
        for (int c = 0; c < dim; ++c) vu[c] = vd[c] = params[c];
        vu[col] = fmin(params[col] + BUMP[col], HI[col]);
        vd[col] = fmax(params[col] - BUMP[col], LO[col]);
        const double h = vu[col] - vd[col];
        if (h <= 0.0) continue;
        for (int64_t i = 0; i < n; ++i) {
            const double ru = single_residual(quotes + 8 * i, model, vu);
            const double rd = single_residual(quotes + 8 * i, model, vd);
            jac[i * dim + col] = (ru - rd) / h;
        }
    }
}

/* calibration.rs:655-667 */
void ora_calib_jtj_grad(const double *jac, const double *res, const double *w, int64_t n, int dim, double *jtj,
                        double *grad) {
    for (int a = 0; a < dim; ++a) {
        grad[a] = 0.0;
        for (int b = 0; b < dim; ++b) jtj[a * dim + b] = 0.0;
    }
    for (int64_t row = 0; row < n; ++row) {
        const double w2 = w[row] * w[row];
        for (int c1 = 0; c1 < dim; ++c1) {
            grad[c1] += w2 * jac[row * dim + c1] * res[row];
            for (int c2 = 0; c2 < dim; ++c2) jtj[c1 * dim + c2] += w2 * jac[row * dim + c1] * jac[row * dim + c2];
        }
    }
}
