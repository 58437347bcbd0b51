/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference's Monte Carlo engine (src/mc.rs) and of the
 * analytic Heston/Bates/BSM pricers it is tested against (src/heston.rs,
 * src/bates.rs, src/bsm.rs).  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this library, and only as the checker
 * or the timed CPU baseline.  The product path (the HIP library) never links
 * or calls it.
 *
 * Functions, with the reference lines each one restates:
 *   ora_splitmix64            mc.rs:337-343
 *   ora_smallrng_seed         mc.rs:172 (rand 0.8 SmallRng::seed_from_u64, see mc_ref.c)
 *   ora_xoshiro_next          rand 0.8 SmallRng = Xoshiro256++ (64-bit targets)
 *   ora_gen_f64 / ora_gen_range_f64   rand 0.8 Standard / UniformFloat::sample_single
 *   ora_box_muller            mc.rs:312-318
 *   ora_sample_poisson        mc.rs:321-331
 *   ora_ncdf                  math.rs:16-19
 *   ora_qe_variance_step      mc.rs:288-309
 *   ora_run_paths             mc.rs:210-278 (+ Payoff::eval mc.rs:82-95)
 *   ora_draw_steps_from_state draw_step mc.rs:129-145 from a given generator state
 *   ora_gen_draws             mc.rs:159-180 (chunked seeding + draw_step mc.rs:129-145)
 *   ora_mc_price              mc.rs:148-207 (mc_heston 98-103 / mc_bates 107-113)
 *   ora_heston_price          heston.rs:61-229
 *   ora_bates_price           bates.rs:29-83
 *   ora_bates_jump_sens       ad.rs:882-895 (price and d/d lambda, mu_j, sigma_j)
 *   ora_cf_price_many         a chain of the two above (batch.rs:23-42)
 *   ora_cf_greeks             greeks.rs:19-82 bump-and-reprice over the two above
 *   ora_bsm_greeks            bsm.rs:11-45 (and ora_bsm_greeks_many, a chain)
 *   ora_implied_vol           iv.rs:14-117 (and ora_implied_vol_many, a chain)
 *   ora_calib_residuals       calibration.rs:573-587 (calib.c)
 *   ora_calib_jacobian        calibration.rs:613-652, Bates jump columns by AD (calib.c)
 *   ora_calib_jtj_grad        calibration.rs:655-667 (calib.c)
 *   ora_bsm_price             bsm.rs:7-12, 55-60
 *   ora_merton_price          textbook Merton (1976) series (no reference counterpart)
 *   ora_build_flags           the compiler command line the library was built with
 */
#ifndef OPE_ORACLE_H
#define OPE_ORACLE_H
#include <stdint.h>

typedef struct { double v0, kappa, theta, sigma, rho; } ora_heston;   /* types.rs:49-55 */
typedef struct { int32_t kind, opt_type; double strike, barrier, rebate; } ora_payoff; /* mc.rs:67-72 */
typedef struct { uint64_t n_paths, n_steps, seed; int32_t antithetic, scheme; } ora_config; /* mc.rs:49-55 */
typedef struct { int32_t has_jumps; double lambda, mu_j, sigma_j; } ora_jumps;   /* mc.rs:107-113 */

enum { ORA_EUROPEAN = 0, ORA_ASIAN = 1, ORA_UP_AND_OUT = 2 };
enum { ORA_CALL = 0, ORA_PUT = 1 };
enum { ORA_EULER = 0, ORA_QE = 1 };

#endif
