/*
 * ope_calib.h -- the evaluations of one Levenberg-Marquardt calibration step
 * on the GPU: implied-volatility residuals of a quote set under a Heston or
 * Bates parameter vector, their Jacobian (central differences, and for the
 * Bates jump columns the exact price derivative over vega), and J'J / J'r.
 *
 * They replace the private helpers the reference's calibrate_* drivers call
 * every iteration (no public counterpart):
 *   residuals of every quote     <- calibration::residuals     (src/calibration.rs:573-587)
 *   Jacobian, J'J, J'r           <- calibration::jacobian,
 *                                   calibration::jtj_and_grad  (src/calibration.rs:613-667)
 * A residual is iv(model price) - iv_market, or 0 where the implied-vol
 * solver finds no volatility (calibration.rs:573-581).  Column c of the
 * Jacobian is (r(p + h_c e_c) - r(p - h_c e_c)) / (actual distance), with
 * the bump points clamped to the parameter bounds and a zero column when a
 * parameter is pinned at both (calibration.rs:637-650); bump sizes and
 * bounds are the reference's (calibration.rs:123-126, 162-171).  All prices
 * of an evaluation (the base point and two per column) go through one launch
 * of the analytic pricer, and all their implied volatilities through one
 * launch of the solver.
 *
 * Bates jump columns (lambda, mu_j, sigma_j) are the reference's AD columns
 * (calibration.rs:620-636, 186-198): dPrice/dparam of every quote, integrated
 * alongside the price on the price's own adaptive panels (the derivative of
 * the reference's forward-mode-AD quadrature, ad.rs:609-731, 882), divided by
 * the BSM vega at the model's implied vol (calibration.rs:593-605).  When a
 * quote's implied vol is not found or its vega is below 1e-10, the whole
 * column is differenced instead (calibration.rs:622-632).
 *
 * params: the parameter vector in CalibModel::to_vec order (calibration.rs:
 * 105, 145-148): v0, kappa, theta, sigma, rho [, lambda, mu_j, sigma_j].
 */
#ifndef OPE_CALIB_H
#define OPE_CALIB_H

#include "ope_cf.h"

#ifdef __cplusplus
extern "C" {
#endif

/* calibration.rs:38-42 (the contract by value; its vol field is unused) */
typedef struct ope_calib_quote {
    ope_option_contract contract;
    double iv_market;
    double weight;
} ope_calib_quote;

/* model: OPE_CF_HESTON (5 parameters) or OPE_CF_BATES (8) */
int ope_calib_residuals(const ope_calib_quote *quotes, size_t n, int32_t model, const double *params,
                        double *residuals);
/* jacobian: n x dim, row-major.  residuals (n), jtj (dim x dim) and grad
 * (dim) are filled when non-null; jtj and grad weight each row by weight^2. */
int ope_calib_jacobian(const ope_calib_quote *quotes, size_t n, int32_t model, const double *params,
                       double *residuals, double *jacobian, double *jtj, double *grad);

#ifdef __cplusplus
}
#endif

#endif /* OPE_CALIB_H */
