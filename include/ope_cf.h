/*
 * ope_cf.h -- the reference's analytic pricers on the GPU: Heston / Bates
 * European prices by characteristic-function inversion (one wavefront per
 * option), their bump-and-reprice Greeks, Black-Scholes-Merton prices and
 * Greeks and implied volatilities (one lane per option), single and chain.
 *
 * Drop-in replacements (same arguments and meaning; a status code and an
 * output pointer instead of the bare f64):
 *   ope_heston_price   <- heston::heston_price   (src/heston.rs:61-72)
 *   ope_bates_price    <- bates::bates_price     (src/bates.rs:29-39)
 *   ope_batch_heston   <- batch::batch_heston    (src/batch.rs:23-27)
 *   ope_batch_bates    <- batch::batch_bates     (src/batch.rs:38-42)
 *   ope_heston_price_and_greeks <- heston::heston_price_and_greeks (src/heston.rs:74-80)
 *   ope_bates_price_and_greeks  <- bates::bates_price_and_greeks   (src/bates.rs:41-47)
 *   ope_batch_heston_greeks     <- batch::batch_heston_greeks      (src/batch.rs:32-36)
 *   ope_batch_bates_greeks      <- batch::batch_bates_greeks       (src/batch.rs:44-48)
 *   ope_bsm_price_and_greeks    <- bsm::bsm_price_and_greeks       (src/bsm.rs:11-35)
 *   ope_bsm_price               <- bsm::bsm_price                  (src/bsm.rs:49-54)
 *   ope_implied_vol             <- iv::implied_vol                 (src/iv.rs:14-20)
 *   ope_batch_bsm               <- batch::batch_bsm                (src/batch.rs:13-15)
 *   ope_batch_bsm_price         <- batch::batch_bsm_price          (src/batch.rs:18-20)
 *   ope_batch_implied_vol       <- batch::batch_implied_vol        (src/batch.rs:51-56)
 * plus ope_cf_price_options, which takes a model per option (what a
 * calibration's residual and Jacobian evaluations need), and the _ex chain
 * variants that also report the GPU time.
 *
 * BSM and implied volatility run one lane per option with the reference's
 * expressions, including its theta, whose carry term has the opposite sign of
 * -dPrice/dT (bsm.rs:38-45).  An implied volatility the reference returns as
 * None comes back as found = 0 and vol = NaN.
 *
 * Greeks are the reference's bump-and-reprice (greeks.rs:19-82): 14 prices
 * per option (spot +-1 %, sqrt(v0) +-1 and +-2 vol points, expiry - 1 day,
 * rate +-1 bp and the cross bumps), all of a batch's bumped prices in one
 * launch, combined on the host with the reference's difference formulas.
 *
 * The quadrature is the reference's: Gil-Pelaez inversion of the Albrecher
 * form CF (heston.rs:82-127, bates.rs:49-83), globally adaptive
 * Gauss-Kronrod 15/7 on u = (1 - t)/t over [0, 1], tolerance 1e-8, at most
 * 200 panels (heston.rs:180-229), with the same bisection decisions, so a
 * price matches the reference's to the last bits of the FP64 special
 * functions.  The reference's own behaviour is kept, including that the CF
 * drift carries r only while the spot is discounted at q (heston.rs:126).
 *
 * Errors: OPE_ERR_INVALID_ARGUMENT for a null pointer, a non-finite
 * contract or model argument, an unknown option type, or (CF pricers)
 * spot / strike <= 0; OPE_ERR_NO_DEVICE without a GPU.  There is no CPU
 * fallback.
 */
#ifndef OPE_CF_H
#define OPE_CF_H

#include "ope_mc.h"

#ifdef __cplusplus
extern "C" {
#endif

/* types.rs:67-72 */
typedef struct ope_bates_params {
    ope_heston_params heston;
    double lambda;  /* jumps per year */
    double mu_j;    /* mean log-jump */
    double sigma_j; /* log-jump volatility */
} ope_bates_params;

/* types.rs:17-25 (vol is the BSM volatility; the CF pricers ignore it) */
typedef struct ope_option_contract {
    double spot;
    double strike;
    double expiry;
    double rate;
    double div_yield;
    double vol;
    int32_t opt_type; /* OPE_CALL / OPE_PUT */
    int32_t reserved_;
} ope_option_contract;

typedef struct ope_cf_option {
    ope_option_contract contract;
    ope_bates_params params; /* model == OPE_CF_HESTON: only params.heston is read */
    int32_t model;           /* OPE_CF_HESTON or OPE_CF_BATES */
    int32_t reserved_;
} ope_cf_option;

enum { OPE_CF_HESTON = 0, OPE_CF_BATES = 1 };

/* types.rs:29-39 */
typedef struct ope_pricing_result {
    double price;
    double delta;
    double gamma;
    double vega;  /* per unit of volatility sqrt(v0) */
    double theta; /* (price(T - 1 day) - price) / 1 day */
    double rho;
    double vanna;
    double volga;
} ope_pricing_result;

/* heston.rs:61-72 */
int ope_heston_price(double spot, double strike, double expiry, double rate, double div_yield,
                     const ope_heston_params *params, int32_t opt_type, double *out);
/* bates.rs:29-39 */
int ope_bates_price(double spot, double strike, double expiry, double rate, double div_yield,
                    const ope_bates_params *params, int32_t opt_type, double *out);
/* batch.rs:23-27: out[i] = heston_price(contracts[i], params), one launch */
int ope_batch_heston(const ope_option_contract *contracts, size_t n, const ope_heston_params *params, double *out);
/* batch.rs:38-42 */
int ope_batch_bates(const ope_option_contract *contracts, size_t n, const ope_bates_params *params, double *out);
/* heston.rs:74-80 */
int ope_heston_price_and_greeks(double spot, double strike, double expiry, double rate, double div_yield,
                                const ope_heston_params *params, int32_t opt_type, ope_pricing_result *out);
/* bates.rs:41-47 */
int ope_bates_price_and_greeks(double spot, double strike, double expiry, double rate, double div_yield,
                               const ope_bates_params *params, int32_t opt_type, ope_pricing_result *out);
/* batch.rs:32-36 */
int ope_batch_heston_greeks(const ope_option_contract *contracts, size_t n, const ope_heston_params *params,
                            ope_pricing_result *out);
/* batch.rs:44-48 */
int ope_batch_bates_greeks(const ope_option_contract *contracts, size_t n, const ope_bates_params *params,
                           ope_pricing_result *out);
/* bsm.rs:11-35 (vol = contract.vol) */
int ope_bsm_price_and_greeks(const ope_option_contract *contract, ope_pricing_result *out);
/* bsm.rs:49-54 */
int ope_bsm_price(const ope_option_contract *contract, double *out);
/* iv.rs:14-20 (contract.vol is ignored) */
int ope_implied_vol(const ope_option_contract *contract, double market_price, double *vol, uint8_t *found);
/* batch.rs:13-15 */
int ope_batch_bsm(const ope_option_contract *contracts, size_t n, ope_pricing_result *out);
/* batch.rs:18-20 */
int ope_batch_bsm_price(const ope_option_contract *contracts, size_t n, double *out);
/* batch.rs:51-56 */
int ope_batch_implied_vol(const ope_option_contract *contracts, const double *market_prices, size_t n, double *vol,
                          uint8_t *found);
/* One launch over a chain: prices (out_price) or prices and Greeks (out_greeks),
 * whichever is non-null (out_greeks wins); kernel_ms, when given, the GPU time. */
int ope_batch_bsm_ex(const ope_option_contract *contracts, size_t n, double *out_price, ope_pricing_result *out_greeks,
                     double *kernel_ms);
int ope_batch_implied_vol_ex(const ope_option_contract *contracts, const double *market_prices, size_t n,
                             double *vol, uint8_t *found, double *kernel_ms);
/* n options with their own models in one launch.  panels, when given,
 * receives 2 entries per option (panels used by the P1 and P2 integrals);
 * kernel_ms, when given, the GPU time of the launch. */
int ope_cf_price_options(const ope_cf_option *options, size_t n, double *out, uint32_t *panels, double *kernel_ms);

#ifdef __cplusplus
}
#endif

#endif /* OPE_CF_H */
