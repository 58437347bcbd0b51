/*
 * ope_mc.h -- C ABI of the MI355X Heston/Bates Monte Carlo engine.
 *
 * Drop-in boundary for the reference's Monte Carlo surface (src/mc.rs,
 * re-exported at src/lib.rs:27).  Each declaration names the reference item it
 * replaces.  All structs are plain C (repr(C) on the Rust side), all inputs
 * are borrowed for the duration of the call only, outputs are caller-owned.
 *
 * Error behaviour: the reference functions are infallible and return NaN on
 * degenerate input (n_paths < 2 with antithetic pairs, QE with sigma = 0 or
 * kappa = 0, mc.rs:160-161, 198-199, 233-235, 293-294).  Here every entry
 * point returns OPE_OK (0) or a negative status and never a silent NaN; a
 * Rust binding panics on a non-zero status to keep the infallible signature.
 * There is no CPU fallback: without a usable GPU the calls return
 * OPE_ERR_NO_DEVICE.
 *
 * Threading: every entry point is blocking and thread-safe (calls on one
 * device are serialised internally).  Results depend only on the inputs:
 * draws are keyed by (seed, pair index), and partial sums are reduced in a
 * fixed order, so they are independent of host threads, launch shape and of
 * other work sharing the GPU.
 */
#ifndef OPE_MC_H
#define OPE_MC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes */
#define OPE_OK 0
#define OPE_ERR_INVALID_ARGUMENT (-1)
#define OPE_ERR_NO_DEVICE (-2)
#define OPE_ERR_HIP (-3)
#define OPE_ERR_OUT_OF_MEMORY (-4)
#define OPE_ERR_OVER_BUDGET (-5)
#define OPE_ERR_NOT_FOUND (-6)
#define OPE_ERR_SHUTDOWN (-7)
#define OPE_ERR_CANCELLED (-8) /* serving: the request was cancelled */
#define OPE_ERR_DEADLINE (-9)  /* serving: the deadline passed before the request was admitted */
#define OPE_ERR_TIMEOUT (-10)  /* serving: the watchdog saw no round complete for watchdog_ms */

/* OptionType (types.rs:5-13) */
#define OPE_CALL 0
#define OPE_PUT 1

/* Payoff variants (mc.rs:66-72) */
#define OPE_PAYOFF_EUROPEAN 0
#define OPE_PAYOFF_ASIAN_ARITHMETIC 1
#define OPE_PAYOFF_UP_AND_OUT 2

/* VarianceScheme (mc.rs:34-46) */
#define OPE_SCHEME_FULL_TRUNCATION_EULER 0
#define OPE_SCHEME_QUADRATIC_EXPONENTIAL 1

/* Extension knobs (no reference counterpart) */
#define OPE_PRECISION_FP64 0      /* all path arithmetic in FP64 (default) */
#define OPE_PRECISION_FP32_PATH 1 /* FP32 path math, FP64 log-price sum, setup and reduction */
#define OPE_RNG_PHILOX 0          /* Philox4x32-10 substreams keyed by (seed, pair) (default) */
#define OPE_RNG_REPLAY 1          /* caller-supplied draws (correctness mode) */

/* HestonParams (types.rs:47-55).  v0 is a variance, not a vol. */
typedef struct ope_heston_params {
    double v0;
    double kappa;
    double theta;
    double sigma;
    double rho;
} ope_heston_params;

/* Payoff (mc.rs:66-72).  The Rust enum becomes a tag plus the union of its
 * fields: barrier and rebate are read only for OPE_PAYOFF_UP_AND_OUT.
 * opt_type is OPE_CALL or OPE_PUT. */
typedef struct ope_payoff {
    int32_t kind;
    int32_t opt_type;
    double strike;
    double barrier;
    double rebate;
} ope_payoff;

/* McConfig (mc.rs:48-55); usize becomes uint64_t.  The reference default
 * (mc.rs:57-64) is {200000, 252, 0xC0FFEE, 1, OPE_SCHEME_FULL_TRUNCATION_EULER},
 * see ope_mc_config_default. */
typedef struct ope_mc_config {
    uint64_t n_paths;
    uint64_t n_steps;
    uint64_t seed;
    uint8_t antithetic;
    uint8_t reserved_[3];
    int32_t scheme;
} ope_mc_config;

/* McResult (mc.rs:28-32) */
typedef struct ope_mc_result {
    double price;
    double std_error;
} ope_mc_result;

/* Raw accumulators of one priced payoff: the reference's (sum, sum_sq, count)
 * triple (mc.rs:164-196), from which price = sum/count and std_error =
 * sqrt(max(sum_sq/count - price^2, 0)/count) (mc.rs:198-204). */
typedef struct ope_mc_stats {
    double sum;
    double sum_sq;
    uint64_t count;
    uint64_t path_steps; /* simulated path-steps (both antithetic siblings count) */
} ope_mc_stats;

/* Replay draws: the three per-pair-step values of StepDraw (mc.rs:126-127),
 * host arrays laid out [n_steps][n_pairs] (pair index fastest).  jump_sum may
 * be NULL for Heston.  n_pairs must equal n_paths/2 (antithetic) or n_paths. */
typedef struct ope_replay_draws {
    const double *za;
    const double *zb;
    const double *jump_sum;
    uint64_t n_pairs;
    uint64_t n_steps;
} ope_replay_draws;

/* Extension options for the *_ex / surface / batch entry points.  A NULL
 * options pointer means {FP64, Philox}. */
typedef struct ope_mc_options {
    int32_t precision;              /* OPE_PRECISION_* */
    int32_t rng;                    /* OPE_RNG_* */
    const ope_replay_draws *replay; /* required when rng == OPE_RNG_REPLAY */
    double *path_payoffs;           /* optional host out, replay only: [n_pairs][2] discounted payoffs */
    double *path_states;            /* optional host out, replay only: [n_pairs][2][4] = (ln S_T, v_T,
                                       running_sum, running_max); untracked entries are NaN */
    ope_mc_stats *stats;            /* optional host out: one per priced result */
    double *device_ms;              /* optional host out: GPU time of the launch (pricing + reduction) */
    double *kernel_ms;              /* optional host out: GPU time of the path-simulation kernel alone */
    uint32_t flags;                 /* OPE_FLAG_* */
    uint32_t reserved_;
} ope_mc_options;

/* FP32-path up-and-out: a path whose running maximum lands within 2^-16 (in
 * log2(S/S0)) of the barrier is simulated again by the FP64 path on the same
 * draws and takes that path's knock-out decision and terminal value, so the
 * decision near B is FP64's.  Off by default: the FP64 pass of a few rare
 * paths extends the kernel's critical path (measured +0.8 ms on the 1M x 365
 * QE up-and-out call, whose FP32 price is already within 0.003 SE of FP64). */
#define OPE_FLAG_BARRIER_FP64 1u

/* Writes the reference's McConfig::default() (mc.rs:57-64). */
void ope_mc_config_default(ope_mc_config *cfg);

/* mc_heston (mc.rs:97-103): Heston path simulation, FP64, Philox draws. */
int ope_mc_heston(double spot, double expiry, double rate, double div_yield,
                  const ope_heston_params *params, const ope_payoff *payoff,
                  const ope_mc_config *cfg, ope_mc_result *out);

/* mc_bates (mc.rs:105-113): Heston plus lognormal jumps with an exact
 * per-step Poisson count.  Takes HestonParams plus the three jump scalars,
 * as the reference does (not BatesParams).  lambda == 0 gives exactly
 * ope_mc_heston's result, as in the reference. */
int ope_mc_bates(double spot, double expiry, double rate, double div_yield,
                 const ope_heston_params *params, double lambda, double mu_j, double sigma_j,
                 const ope_payoff *payoff, const ope_mc_config *cfg, ope_mc_result *out);

/* mc_heston / mc_bates with extension options (precision, replay draws,
 * per-path outputs, statistics).  No reference counterpart beyond mc.rs:98-113. */
int ope_mc_heston_ex(double spot, double expiry, double rate, double div_yield,
                     const ope_heston_params *params, const ope_payoff *payoff,
                     const ope_mc_config *cfg, const ope_mc_options *opts, ope_mc_result *out);
int ope_mc_bates_ex(double spot, double expiry, double rate, double div_yield,
                    const ope_heston_params *params, double lambda, double mu_j, double sigma_j,
                    const ope_payoff *payoff, const ope_mc_config *cfg,
                    const ope_mc_options *opts, ope_mc_result *out);

/* No counterpart in the reference.  ope_mc_bates_ex for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the call allocates nothing; its device buffers live in a
 * fixed 256.25 KiB array in the library's device code (one per GPU, held
 * until the process ends).  The first call on a device also initialises the
 * library's context and looks the array up.  Philox draws only (no replay,
 * no per-path outputs); at most 2^24 antithetic pairs (2^25 paths), larger
 * requests fail with OPE_ERR_INVALID_ARGUMENT.  The result equals
 * ope_mc_bates_ex's bit for bit (lambda == 0: ope_mc_heston's).  Calls are
 * serialised by the library; the array keeps the last call's contents. */
int ope_mc_bates_noalloc(double spot, double expiry, double rate, double div_yield,
                         const ope_heston_params *params, double lambda, double mu_j, double sigma_j,
                         const ope_payoff *payoff, const ope_mc_config *cfg,
                         const ope_mc_options *opts, ope_mc_result *out);

/* European strike x expiry surface from shared paths.  One simulation per
 * expiry (dt = expiry / cfg->n_steps, the reference's semantics), every
 * strike priced from the same paths.  out[e * n_strikes + k] equals what
 * ope_mc_bates(spot, expiries[e], ..., European(strikes[k]), cfg) returns
 * (the reference would re-simulate the same seeded paths once per strike:
 * 328 mc_bates calls for 41 strikes x 8 expiries).  lambda == 0 is Heston.
 * opts->stats, when given, receives n_expiries * n_strikes entries. */
int ope_mc_surface(double spot, double rate, double div_yield, const ope_heston_params *params,
                   double lambda, double mu_j, double sigma_j, int32_t opt_type,
                   const double *expiries, uint32_t n_expiries, const double *strikes,
                   uint32_t n_strikes, const ope_mc_config *cfg, const ope_mc_options *opts,
                   ope_mc_result *out);

/* One independent pricing request: the full argument list of mc_bates
 * (lambda == 0 means mc_heston). */
typedef struct ope_mc_request {
    double spot;
    double expiry;
    double rate;
    double div_yield;
    ope_heston_params heston;
    double lambda;
    double mu_j;
    double sigma_j;
    ope_payoff payoff;
    ope_mc_config cfg;
} ope_mc_request;

/* Uneven batch in ONE kernel launch: requests with different models, payoffs,
 * path and step counts.  out[i] is bit-identical to pricing request i alone
 * with ope_mc_bates_ex and the same options.  Replay is not accepted here.
 * opts->stats, when given, receives n entries. */
int ope_mc_batch(const ope_mc_request *reqs, uint32_t n, const ope_mc_options *opts,
                 ope_mc_result *out);

/* ---- serving: persistent-kernel request loop (no reference counterpart)
 *
 * Requests are admitted against a device-memory budget and a number of
 * resident slots, in arrival order.  A persistent kernel keeps running the
 * tiles of every admitted request; requests that arrive while others are on
 * the GPU join as soon as a slot and budget free up, and their tiles run
 * alongside the others'.  Each request's pair counter and partial sums stay
 * resident on the device between rounds.  A request with target_rel_se > 0
 * runs in rounds of 256-pair tiles until std_error <= target_rel_se * |price|
 * or its n_paths are spent; the decision uses only the request's own sums, so
 * a result does not depend on what else is being served.  A request that
 * stops after P pairs returns bit-for-bit what ope_mc_bates_ex returns for
 * n_paths = 2P (or P without antithetic pairs) when that call runs one pair
 * per thread (fewer than 2^22 pairs). */
typedef struct ope_server ope_server;

typedef struct ope_server_config {
    int32_t precision;            /* OPE_PRECISION_* for every request */
    uint32_t max_requests;        /* resident request slots (0: 1024) */
    uint64_t memory_budget_bytes; /* admission budget for resident request state (0: 64 GiB) */
    uint32_t initial_round_pairs; /* pairs in a request's first round (0: 65536) */
    uint32_t ring_items;          /* work-ring capacity in 256-pair tiles (0: 2^22); the tiles of
                                     every resident request, over all its rounds, must fit */
    double idle_exit_ms;          /* the persistent kernel exits after this long without work (0: 200) */
    double watchdog_ms;           /* with tiles outstanding and no round completing for this long,
                                     every outstanding request fails with OPE_ERR_TIMEOUT and the
                                     server stops admitting (0: 30000; negative: off) */
} ope_server_config;

/* Why a served request stopped (ope_server_request_info.stop_reason). */
#define OPE_SERVER_STOP_ALL_PATHS 0 /* ran its n_paths */
#define OPE_SERVER_STOP_TARGET 1    /* std_error <= target_rel_se * |price| */
#define OPE_SERVER_STOP_DEADLINE 2  /* its deadline passed at a round boundary: the result is the
                                       estimate from the paths run so far */

typedef struct ope_server_request_info {
    uint64_t paths;        /* paths simulated (both antithetic siblings count) */
    uint64_t path_steps;
    uint32_t rounds;
    uint32_t slot;
    double queue_ms;       /* submit -> admission */
    double latency_ms;     /* submit -> result */
    int32_t stop_reason;   /* OPE_SERVER_STOP_* */
    uint32_t pad_;
} ope_server_request_info;

typedef struct ope_server_stats {
    uint64_t submitted;
    uint64_t completed;
    uint64_t path_steps;          /* simulated by completed requests */
    uint64_t peak_resident_bytes; /* largest admitted footprint at one time */
    uint32_t peak_active;         /* most requests resident at one time */
    uint32_t kernel_launches;     /* persistent-kernel launches (relaunched after idle exits) */
    uint32_t workgroups;          /* workgroups per launch (one of them polls the host) */
    uint32_t pad_;
} ope_server_stats;

int ope_server_create(const ope_server_config *cfg, ope_server **out);
/* Queues one request (validated here).  target_rel_se <= 0: run all n_paths. */
int ope_server_submit(ope_server *s, const ope_mc_request *req, double target_rel_se, uint64_t *ticket);
/* As ope_server_submit, with a deadline deadline_ms after submission (<= 0:
 * none).  A request still queued at its deadline fails with OPE_ERR_DEADLINE;
 * an admitted request stops at the first round boundary past its deadline and
 * returns the estimate from the paths it ran (stop_reason
 * OPE_SERVER_STOP_DEADLINE; a request without a target runs in one round, so
 * the deadline bounds only its wait for admission). */
int ope_server_submit_ex(ope_server *s, const ope_mc_request *req, double target_rel_se, double deadline_ms,
                         uint64_t *ticket);
/* Cancels a request: a queued one fails at once, an admitted one at the end of
 * the round it is running (its posted tiles still run).  Waiting on it then
 * returns OPE_ERR_CANCELLED.  OPE_ERR_NOT_FOUND if the ticket is unknown or
 * has already finished. */
int ope_server_cancel(ope_server *s, uint64_t ticket);
/* *done = 1 and the outputs filled once the request has finished. */
int ope_server_poll(ope_server *s, uint64_t ticket, ope_mc_result *out, ope_server_request_info *info,
                    int32_t *done);
/* Blocks until the request finishes or timeout_ms passes (OPE_ERR_NOT_FOUND if
 * it has not finished; negative timeout: no limit).  A finished ticket is
 * forgotten after a successful wait. */
int ope_server_wait(ope_server *s, uint64_t ticket, double timeout_ms, ope_mc_result *out,
                    ope_server_request_info *info);
int ope_server_get_stats(ope_server *s, ope_server_stats *out);
/* Finishes the admitted requests, fails the queued ones, stops the kernel. */
int ope_server_destroy(ope_server *s);

/* ---- device / diagnostics */
typedef struct ope_device_info {
    char name[64];
    char arch[32];
    int32_t compute_units;
    int32_t clock_mhz;
    uint64_t total_mem_bytes;
} ope_device_info;

int ope_get_device_info(ope_device_info *info);

/* Human-readable text of a status code (static storage). */
const char *ope_status_string(int status);

/* Detail of the last failure on the calling thread (empty if none). */
const char *ope_last_error_message(void);

/* Philox4x32-10 on the device: ctr[4*i..4*i+3] with key (seed lo, seed hi)
 * -> out[4*i..4*i+3].  Exposed so the generator can be checked against
 * known-answer vectors. */
int ope_debug_philox(uint64_t seed, const uint32_t *ctr, uint32_t n, uint32_t *out);

/* The FP64 Philox-mode path's elementary functions on the device, for
 * accuracy checks (no reference counterpart): fn 0 natural log, 1 exp,
 * 2 / 3 sine / cosine of 2 pi x 2^-23 for integer x in [0, 2^23) (the
 * Box-Muller angle), 4 reciprocal, 5 reciprocal square root, 6 square
 * root (x = 0 or a positive normal), 7 normal tail Phi(-|x|) for
 * |x| <= 6.8; out[i] = f(x[i]). */
int ope_debug_fp64_math(int32_t fn, const double *x, uint32_t n, double *out);

/* The Philox-mode draws of pairs [pair0, pair0+n_pairs) over n_steps steps as
 * the pricing kernels see them, written [n_steps][n_pairs] as doubles (FP32
 * draws widened).  lambda > 0 selects the Bates stream (lambda * expiry /
 * n_steps is the Poisson mean); n_jumps (optional) receives the Poisson
 * counts. */
int ope_debug_draws(uint64_t seed, uint64_t pair0, uint64_t n_pairs, uint64_t n_steps,
                    double expiry, double lambda, double mu_j, double sigma_j, int32_t precision,
                    double *za, double *zb, double *jump_sum, uint32_t *n_jumps);

#ifdef __cplusplus
}
#endif

#endif /* OPE_MC_H */
