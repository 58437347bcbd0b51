/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * C restatement of the reference CPU Monte Carlo engine, src/mc.rs, compiled
 * with -ffp-contract=off so that no a*b+c is fused (rustc never fuses), and
 * against glibc libm, which is what Rust's f64::ln/exp/sqrt/sin/cos call on
 * Linux.  Every arithmetic expression keeps the reference's evaluation order
 * (Rust evaluates left to right), so that given the same draws the per-path
 * state agrees bit for bit with the reference's arithmetic.
 *
 * Third-party semantics restated here (the crates are absent from the
 * reference tree and unpinned in its Cargo.lock):
 *   - rand 0.8.5 SmallRng on 64-bit targets is Xoshiro256++ (Blackman-Vigna).
 *   - SmallRng::seed_from_u64 in rand 0.8 does not forward to the generator's
 *     own SplitMix64 seeding; it falls through to rand_core 0.6's default,
 *     which fills the 32-byte seed with PCG32 (XSH-RR) outputs, advancing the
 *     LCG state before each output.  Pinned by rand_core 0.6's own
 *     value-breakage answer (seed 0 expands to first word
 *     5029875928683246316, tests/test_oracle_rng.py); the uniform maps and
 *     draw_step's word order below are pinned there too, so crate-identical
 *     draws are open only up to the two last-bit questions that follow.
 *   - gen::<f64>() = (next_u64 >> 11) * 2^-53.
 *   - gen_range(lo..hi) for f64 = UniformFloat::sample_single: build a value in
 *     [1,2) from the top 52 bits, subtract 1, scale, add lo; retry if >= hi.
 *   - libm::erfc (musl/fdlibm algorithm) is replaced by glibc erfc, which is
 *     the same fdlibm algorithm; last-bit agreement is unverified.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ RNG */

/* mc.rs:337-343 */
uint64_t ora_splitmix64(uint64_t x) {
    x = x + 0x9E3779B97F4A7C15ULL;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

typedef struct { uint64_t s[4]; } xoshiro256;

static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

/* Xoshiro256++ next(): output rotl(s0 + s3, 23) + s0, then the xoshiro256
 * state transition. */
static inline uint64_t xoshiro_next(xoshiro256 *g) {
    uint64_t *s = g->s;
    const uint64_t result = rotl64(s[0] + s[3], 23) + s[0];
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return result;
}

/* PCG32 XSH-RR output of a 64-bit LCG state. */
static inline uint32_t pcg_xsh_rr(uint64_t state) {
    uint32_t xorshifted = (uint32_t)(((state >> 18) ^ state) >> 27);
    uint32_t rot = (uint32_t)(state >> 59);
    return (xorshifted >> rot) | (xorshifted << ((32 - rot) & 31));
}

/* rand_core 0.6 SeedableRng::seed_from_u64 default: for each 4-byte chunk of
 * the seed, advance state = state*MUL + INC, then emit the XSH-RR output of
 * the new state, little endian.  The 32 bytes are then read as four
 * little-endian u64 words (Xoshiro256++::from_seed).  An all-zero seed never
 * arises from this expansion in practice; from_seed's zero-seed fallback is
 * restated anyway. */
static void smallrng_seed_from_u64(xoshiro256 *g, uint64_t state) {
    const uint64_t MUL = 6364136223846793005ULL;
    const uint64_t INC = 11634580027462260723ULL;
    uint32_t w[8];
    for (int i = 0; i < 8; ++i) {
        state = state * MUL + INC;
        w[i] = pcg_xsh_rr(state);
    }
    for (int i = 0; i < 4; ++i) g->s[i] = (uint64_t)w[2 * i] | ((uint64_t)w[2 * i + 1] << 32);
    if (!(g->s[0] | g->s[1] | g->s[2] | g->s[3])) {
        /* Xoshiro256PlusPlus::from_seed(all zero) -> seed_from_u64(0) of the
         * generator itself, which is SplitMix64-based. */
        uint64_t x = 0;
        for (int i = 0; i < 4; ++i) {
            x += 0x9E3779B97F4A7C15ULL;
            uint64_t z = x;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
            g->s[i] = z ^ (z >> 31);
        }
    }
}

/* rand 0.8 Standard for f64: 53 high bits times 2^-53, in [0,1). */
static inline double gen_f64(xoshiro256 *g) {
    uint64_t v = xoshiro_next(g) >> 11;
    return (double)v * (1.0 / 9007199254740992.0);
}

/* rand 0.8 UniformFloat<f64>::sample_single(low, high). */
static inline double gen_range_f64(xoshiro256 *g, double low, double high) {
    double scale = high - low;
    for (;;) {
        uint64_t bits = (xoshiro_next(g) >> 12) | (1023ULL << 52);
        double value1_2;
        memcpy(&value1_2, &bits, sizeof bits);
        double value0_1 = value1_2 - 1.0;
        double res = value0_1 * scale + low;
        if (res < high) return res;
        /* Only reachable when low + scale rounds up to high: shrink the scale
         * by one ulp and retry (rand's edge-case handling). */
        scale = nextafter(scale, 0.0);
    }
}

/* mc.rs:312-318 */
static inline void box_muller(xoshiro256 *g, double *z0, double *z1) {
    double u1 = gen_range_f64(g, 1e-12, 1.0);
    double u2 = gen_f64(g);
    double r = sqrt(-2.0 * log(u1));
    double theta = 6.283185307179586 * u2; /* std::f64::consts::TAU */
    *z0 = r * cos(theta);
    *z1 = r * sin(theta);
}

/* mc.rs:321-331 (Knuth) */
static inline uint32_t sample_poisson(xoshiro256 *g, double mean) {
    if (mean <= 0.0) return 0;
    double l = exp(-mean);
    uint32_t k = 0;
    double p = 1.0;
    for (;;) {
        k += 1;
        p *= gen_f64(g);
        if (p <= l) return k - 1;
    }
}

typedef struct { double za, zb, jump_sum; } step_draw; /* mc.rs:126-127 */

/* mc.rs:129-145.  *count (optional) receives the Poisson count. */
static inline step_draw draw_step_n(xoshiro256 *g, const ora_jumps *j, double dt, uint32_t *count) {
    step_draw d;
    box_muller(g, &d.za, &d.zb);
    d.jump_sum = 0.0;
    if (count) *count = 0;
    if (j && j->has_jumps) {
        uint32_t n = sample_poisson(g, j->lambda * dt);
        if (count) *count = n;
        if (n != 0) {
            double u, unused;
            box_muller(g, &u, &unused);
            d.jump_sum = j->mu_j * (double)n + j->sigma_j * sqrt((double)n) * u;
        }
    }
    return d;
}

static inline step_draw draw_step(xoshiro256 *g, const ora_jumps *j, double dt) {
    return draw_step_n(g, j, dt, NULL);
}

/* ------------------------------------------------------------ path math */

#define SQRT_2_INV 0.7071067811865476 /* std::f64::consts::FRAC_1_SQRT_2 */

/* math.rs:16-19 */
double ora_ncdf(double x) { return 0.5 * erfc(-x * SQRT_2_INV); }

static inline double fmax_rs(double a, double b) { return fmax(a, b); } /* f64::max */

/* mc.rs:288-309 */
double ora_qe_variance_step(double v_t, double kappa, double theta, double sigma, double dt, double z) {
    const double PSI_C = 1.5;
    double ekt = exp(-kappa * dt);
    double m = theta + (v_t - theta) * ekt;
    double one_m = 1.0 - ekt;
    double s2 = (v_t * sigma * sigma * ekt / kappa) * one_m
              + (theta * sigma * sigma / (2.0 * kappa)) * (one_m * one_m);
    double psi = fmax_rs(s2 / fmax_rs(m * m, 1e-300), 1e-12);
    if (psi <= PSI_C) {
        double inv_psi = 1.0 / psi;
        double b2 = 2.0 * inv_psi - 1.0 + sqrt(fmax_rs(2.0 * inv_psi * (2.0 * inv_psi - 1.0), 0.0));
        double a = m / (1.0 + b2);
        double b = sqrt(b2);
        double bz = b + z;
        return a * (bz * bz);
    } else {
        double p = (psi - 1.0) / (psi + 1.0);
        double beta = (1.0 - p) / fmax_rs(m, 1e-300);
        double u = ora_ncdf(z);
        if (u <= p) return 0.0;
        return (1.0 / beta) * log((1.0 - p) / (1.0 - u));
    }
}

/* mc.rs:75-80 */
static inline double intrinsic(double s, double k, int opt_type) {
    return opt_type == ORA_CALL ? fmax_rs(s - k, 0.0) : fmax_rs(k - s, 0.0);
}

/* mc.rs:84-95 */
static inline double payoff_eval(const ora_payoff *p, double terminal, double running_sum,
                                 double running_max, uint64_t n_steps) {
    switch (p->kind) {
    case ORA_EUROPEAN: return intrinsic(terminal, p->strike, p->opt_type);
    case ORA_ASIAN: return intrinsic(running_sum / (double)n_steps, p->strike, p->opt_type);
    default:
        if (running_max >= p->barrier) return p->rebate;
        return intrinsic(terminal, p->strike, p->opt_type);
    }
}

typedef struct {
    double spot, drift, dt, sqdt, rate, expiry;
    ora_heston hp;
    int scheme;
    ora_payoff payoff;
} path_ctx;

/* mc.rs:210-278.  state (optional) receives {ln_s, v, running_sum, running_max}. */
static double run_path(const path_ctx *c, const step_draw *steps, uint64_t n_steps, int antithetic,
                       double *state) {
    const ora_heston *hp = &c->hp;
    double ln_s = log(c->spot);
    double v = hp->v0;
    double running_sum = 0.0;
    double running_max = c->spot;
    const double dt = c->dt, sqdt = c->sqdt, drift = c->drift;

    double k0 = 0, k1 = 0, k2 = 0, k3 = 0, k4 = 0;
    if (c->scheme == ORA_QE) { /* mc.rs:230-238, gamma1 = gamma2 = 0.5 */
        double rho = hp->rho, g1 = 0.5, g2 = 0.5;
        k0 = -rho * hp->kappa * hp->theta * dt / hp->sigma;
        k1 = g1 * dt * (hp->kappa * rho / hp->sigma - 0.5) - rho / hp->sigma;
        k2 = g2 * dt * (hp->kappa * rho / hp->sigma - 0.5) + rho / hp->sigma;
        k3 = g1 * dt * (1.0 - rho * rho);
        k4 = g2 * dt * (1.0 - rho * rho);
    }

    for (uint64_t i = 0; i < n_steps; ++i) {
        const step_draw *st = &steps[i];
        double za = antithetic ? -st->za : st->za;
        double zb = antithetic ? -st->zb : st->zb;
        double v_pos = fmax_rs(v, 0.0);
        if (c->scheme == ORA_EULER) { /* mc.rs:248-255 */
            double z1 = za;
            double z2 = hp->rho * za + sqrt(1.0 - hp->rho * hp->rho) * zb;
            ln_s += (drift - 0.5 * v_pos) * dt + sqrt(v_pos) * sqdt * z1 + st->jump_sum;
            v = fmax_rs(v + hp->kappa * (hp->theta - v_pos) * dt + hp->sigma * sqrt(v_pos) * sqdt * z2, 0.0);
        } else { /* mc.rs:256-267 */
            double v_new = ora_qe_variance_step(v_pos, hp->kappa, hp->theta, hp->sigma, dt, za);
            double var_term = fmax_rs(k3 * v_pos + k4 * v_new, 0.0);
            ln_s += drift * dt + k0 + k1 * v_pos + k2 * v_new + sqrt(var_term) * zb + st->jump_sum;
            v = v_new;
        }
        double s = exp(ln_s); /* mc.rs:270-272 */
        running_sum += s;
        running_max = fmax_rs(running_max, s);
    }
    double terminal = exp(ln_s);
    double disc = exp(-c->rate * c->expiry);
    if (state) {
        state[0] = ln_s;
        state[1] = v;
        state[2] = running_sum;
        state[3] = running_max;
    }
    return disc * payoff_eval(&c->payoff, terminal, running_sum, running_max, n_steps);
}

/* mc.rs:153-157 */
static void make_ctx(path_ctx *c, double spot, double expiry, double rate, double div_yield,
                     const ora_heston *hp, const ora_jumps *j, const ora_payoff *payoff,
                     uint64_t n_steps, int scheme) {
    c->spot = spot;
    c->rate = rate;
    c->expiry = expiry;
    c->hp = *hp;
    c->scheme = scheme;
    c->payoff = *payoff;
    c->dt = expiry / (double)n_steps;
    c->sqdt = sqrt(c->dt);
    int jumps = j && j->has_jumps;
    double k_bar = jumps ? exp(j->mu_j + 0.5 * j->sigma_j * j->sigma_j) - 1.0 : 0.0;
    double lambda = jumps ? j->lambda : 0.0;
    c->drift = rate - div_yield - lambda * k_bar;
}

/* Chunk bounds of mc.rs:159-170 for a given thread count. */
static void chunk_bounds(uint64_t n_pairs_total, uint64_t n_threads, uint64_t chunk_idx,
                         uint64_t *start, uint64_t *end) {
    uint64_t chunk_size = n_pairs_total / n_threads;
    if (chunk_size < 1) chunk_size = 1;
    uint64_t s = chunk_idx * chunk_size;
    uint64_t e;
    if (chunk_idx == n_threads - 1) e = n_pairs_total;
    else { e = s + chunk_size; if (e > n_pairs_total) e = n_pairs_total; }
    *start = s;
    *end = e;
}

/* ------------------------------------------------------------ mc_price */

typedef struct {
    const path_ctx *ctx;
    const ora_jumps *jumps;
    const ora_config *cfg;
    uint64_t n_pairs_total, n_threads, chunk_idx;
    double sum, sum_sq;
    uint64_t count;
} chunk_job;

static void *run_chunk(void *arg) {
    chunk_job *jb = (chunk_job *)arg;
    uint64_t start, end;
    chunk_bounds(jb->n_pairs_total, jb->n_threads, jb->chunk_idx, &start, &end);
    jb->sum = 0.0;
    jb->sum_sq = 0.0;
    jb->count = 0;
    if (start >= end) return NULL;
    xoshiro256 g;
    smallrng_seed_from_u64(&g, ora_splitmix64(jb->cfg->seed ^ jb->chunk_idx));
    uint64_t n_steps = jb->cfg->n_steps;
    double local_sum = 0.0, local_sum_sq = 0.0;
    uint64_t local_count = 0;
    for (uint64_t p = start; p < end; ++p) {
        /* mc.rs:178-180: one heap buffer of draws per pair */
        step_draw *steps = (step_draw *)malloc(sizeof(step_draw) * (n_steps ? n_steps : 1));
        for (uint64_t i = 0; i < n_steps; ++i) steps[i] = draw_step(&g, jb->jumps, jb->ctx->dt);
        double pm = run_path(jb->ctx, steps, n_steps, 0, NULL);
        local_sum += pm;
        local_sum_sq += pm * pm;
        local_count += 1;
        if (jb->cfg->antithetic) {
            double pa = run_path(jb->ctx, steps, n_steps, 1, NULL);
            local_sum += pa;
            local_sum_sq += pa * pa;
            local_count += 1;
        }
        free(steps);
    }
    jb->sum = local_sum;
    jb->sum_sq = local_sum_sq;
    jb->count = local_count;
    return NULL;
}

/* mc.rs:148-207.  n_threads plays the role of rayon::current_num_threads():
 * it fixes the chunking and hence the seeds.  workers is how many OS threads
 * actually run the chunks (results do not depend on it).  out[0..4] =
 * {price, std_error, sum, sum_sq}; *count receives the path count. */
int ora_mc_price(double spot, double expiry, double rate, double div_yield, const ora_heston *hp,
                 const ora_jumps *jumps, const ora_payoff *payoff, const ora_config *cfg,
                 uint64_t n_threads, int workers, double *out, uint64_t *count_out) {
    if (n_threads < 1) n_threads = 1;
    path_ctx ctx;
    make_ctx(&ctx, spot, expiry, rate, div_yield, hp, jumps, payoff, cfg->n_steps, cfg->scheme);
    uint64_t paths_per_pair = cfg->antithetic ? 2 : 1;
    uint64_t n_pairs_total = cfg->n_paths / paths_per_pair;

    chunk_job *jobs = (chunk_job *)calloc(n_threads, sizeof(chunk_job));
    if (!jobs) return -1;
    for (uint64_t c = 0; c < n_threads; ++c) {
        jobs[c].ctx = &ctx;
        jobs[c].jumps = jumps;
        jobs[c].cfg = cfg;
        jobs[c].n_pairs_total = n_pairs_total;
        jobs[c].n_threads = n_threads;
        jobs[c].chunk_idx = c;
    }
    if (workers <= 1) {
        for (uint64_t c = 0; c < n_threads; ++c) run_chunk(&jobs[c]);
    } else {
        /* run chunks in waves of `workers` threads */
        pthread_t *th = (pthread_t *)calloc((size_t)workers, sizeof(pthread_t));
        for (uint64_t base = 0; base < n_threads; base += (uint64_t)workers) {
            uint64_t lim = base + (uint64_t)workers < n_threads ? base + (uint64_t)workers : n_threads;
            for (uint64_t c = base; c < lim; ++c) pthread_create(&th[c - base], NULL, run_chunk, &jobs[c]);
            for (uint64_t c = base; c < lim; ++c) pthread_join(th[c - base], NULL);
        }
        free(th);
    }
    double sum = 0.0, sum_sq = 0.0;
    uint64_t count = 0;
    for (uint64_t c = 0; c < n_threads; ++c) {
        sum += jobs[c].sum;
        sum_sq += jobs[c].sum_sq;
        count += jobs[c].count;
    }
    free(jobs);
    double n = (double)count;
    double mean = sum / n;
    double var = fmax_rs(sum_sq / n - mean * mean, 0.0);
    out[0] = mean;
    out[1] = sqrt(var / n);
    out[2] = sum;
    out[3] = sum_sq;
    if (count_out) *count_out = count;
    return 0;
}

/* ------------------------------------------------------- replay helpers */

/* The draws mc_price would consume, for every pair in global pair order,
 * written [step][pair] (za, zb, jump_sum).  Pair p of chunk c is the
 * (p - start_c)-th pair drawn from chunk c's generator. */
int ora_gen_draws(uint64_t seed, uint64_t n_pairs_total, uint64_t n_steps, uint64_t n_threads,
                  double expiry, const ora_jumps *jumps, double *za, double *zb, double *jump) {
    if (n_threads < 1) n_threads = 1;
    double dt = expiry / (double)n_steps;
    for (uint64_t c = 0; c < n_threads; ++c) {
        uint64_t start, end;
        chunk_bounds(n_pairs_total, n_threads, c, &start, &end);
        if (start >= end) continue;
        xoshiro256 g;
        smallrng_seed_from_u64(&g, ora_splitmix64(seed ^ c));
        for (uint64_t p = start; p < end; ++p) {
            for (uint64_t i = 0; i < n_steps; ++i) {
                step_draw d = draw_step(&g, jumps, dt);
                za[i * n_pairs_total + p] = d.za;
                zb[i * n_pairs_total + p] = d.zb;
                jump[i * n_pairs_total + p] = d.jump_sum;
            }
        }
    }
    return 0;
}

/* run_path for every pair on given draws ([step][pair] layout).  Outputs:
 * payoff[pair*2 + sib], state[(pair*2 + sib)*4 + {ln_s, v, sum, max}];
 * sib 1 is written only when antithetic. */
int ora_run_paths(double spot, double expiry, double rate, double div_yield, const ora_heston *hp,
                  const ora_jumps *jumps, const ora_payoff *payoff, uint64_t n_steps, int scheme,
                  int antithetic, uint64_t n_pairs, const double *za, const double *zb,
                  const double *jump, double *payoff_out, double *state_out) {
    path_ctx ctx;
    make_ctx(&ctx, spot, expiry, rate, div_yield, hp, jumps, payoff, n_steps, scheme);
    step_draw *steps = (step_draw *)malloc(sizeof(step_draw) * (n_steps ? n_steps : 1));
    if (!steps) return -1;
    for (uint64_t p = 0; p < n_pairs; ++p) {
        for (uint64_t i = 0; i < n_steps; ++i) {
            steps[i].za = za[i * n_pairs + p];
            steps[i].zb = zb[i * n_pairs + p];
            steps[i].jump_sum = jump ? jump[i * n_pairs + p] : 0.0;
        }
        payoff_out[p * 2] = run_path(&ctx, steps, n_steps, 0, state_out ? state_out + (p * 2) * 4 : NULL);
        if (antithetic)
            payoff_out[p * 2 + 1] = run_path(&ctx, steps, n_steps, 1, state_out ? state_out + (p * 2 + 1) * 4 : NULL);
    }
    free(steps);
    return 0;
}

/* ------------------------------------------------ primitives for tests */

#ifndef ORA_BUILD_FLAGS
#define ORA_BUILD_FLAGS "unknown"
#endif
/* The compiler and flags this library was built with. */
const char *ora_build_flags(void) { return ORA_BUILD_FLAGS; }

void ora_smallrng_seed(uint64_t seed, uint64_t state_out[4]) {
    xoshiro256 g;
    smallrng_seed_from_u64(&g, seed);
    memcpy(state_out, g.s, sizeof g.s);
}

void ora_xoshiro_next_n(uint64_t state[4], uint64_t n, uint64_t *out) {
    xoshiro256 g;
    memcpy(g.s, state, sizeof g.s);
    for (uint64_t i = 0; i < n; ++i) out[i] = xoshiro_next(&g);
    memcpy(state, g.s, sizeof g.s);
}

/* Canonical PCG32 (pcg32_srandom_r / pcg32_random_r) for a known-answer check
 * of the XSH-RR output permutation used by the seeding above. */
void ora_pcg32_canonical(uint64_t initstate, uint64_t initseq, uint64_t n, uint32_t *out) {
    const uint64_t MUL = 6364136223846793005ULL;
    uint64_t inc = (initseq << 1u) | 1u;
    uint64_t state = 0;
    state = state * MUL + inc;
    state += initstate;
    state = state * MUL + inc;
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t old = state;
        state = old * MUL + inc;
        out[i] = pcg_xsh_rr(old);
    }
}

/* n_steps successive draw_step calls (mc.rs:129-145) from a given Xoshiro256++
 * state, which is advanced in place: za, zb, jump_sum and the Poisson count of
 * each step.  Lets a test check which generator outputs each draw consumes. */
void ora_draw_steps_from_state(uint64_t state[4], const ora_jumps *j, double dt, uint64_t n_steps,
                               double *za, double *zb, double *jump_sum, uint32_t *count) {
    xoshiro256 g;
    memcpy(g.s, state, sizeof g.s);
    for (uint64_t i = 0; i < n_steps; ++i) {
        step_draw d = draw_step_n(&g, j, dt, &count[i]);
        za[i] = d.za;
        zb[i] = d.zb;
        jump_sum[i] = d.jump_sum;
    }
    memcpy(state, g.s, sizeof g.s);
}

void ora_gen_f64_n(uint64_t seed, uint64_t n, double *out) {
    xoshiro256 g;
    smallrng_seed_from_u64(&g, seed);
    for (uint64_t i = 0; i < n; ++i) out[i] = gen_f64(&g);
}

void ora_gen_range_n(uint64_t seed, double lo, double hi, uint64_t n, double *out) {
    xoshiro256 g;
    smallrng_seed_from_u64(&g, seed);
    for (uint64_t i = 0; i < n; ++i) out[i] = gen_range_f64(&g, lo, hi);
}

void ora_box_muller_n(uint64_t seed, uint64_t n_pairs, double *out) {
    xoshiro256 g;
    smallrng_seed_from_u64(&g, seed);
    for (uint64_t i = 0; i < n_pairs; ++i) box_muller(&g, &out[2 * i], &out[2 * i + 1]);
}

/* mc.rs:457-465 style: n draws of sample_poisson(mean) from seed_from_u64(seed). */
void ora_sample_poisson_n(uint64_t seed, double mean, uint64_t n, uint32_t *out) {
    xoshiro256 g;
    smallrng_seed_from_u64(&g, seed);
    for (uint64_t i = 0; i < n; ++i) out[i] = sample_poisson(&g, mean);
}

void ora_qe_variance_step_n(const double *v_t, const double *z, uint64_t n, double kappa, double theta,
                            double sigma, double dt, double *out) {
    for (uint64_t i = 0; i < n; ++i) out[i] = ora_qe_variance_step(v_t[i], kappa, theta, sigma, dt, z[i]);
}

void ora_ncdf_n(const double *x, uint64_t n, double *out) {
    for (uint64_t i = 0; i < n; ++i) out[i] = ora_ncdf(x[i]);
}
