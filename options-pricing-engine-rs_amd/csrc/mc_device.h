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

#include <type_traits>

#include "fp64_math.h"
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

// Substreams of one pair (the counter's fourth word):
//   kTagNormal   call s/2: words (x,y) -> normals of step s for even s, (z,w)
//                for odd s; the angle word's 23 high bits give the angle, its
//                9 low bits the 9 high bits of the Poisson uniform of step s
//                (Bates)
//   kTagJump     call s:   word z -> the uniform's 23 low bits, when the high
//                          bits allow a non-zero count; words (x,y) -> the
//                          jump-size normal of step s when the count is non-zero
constexpr uint32_t kTagNormal = 0u, kTagJump = 2u;
constexpr uint32_t kGateBits = 9u, kGateMask = (1u << kGateBits) - 1u;

// a ^ b ^ c in one v_bitop3_b32 (truth table 0x96, the parity of the three
// inputs).  The instruction takes at most one scalar operand, so it is used
// only where the other two are per-lane values.
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

// In the engine's counters (call, pair lo, pair hi, tag) the call index, the
// pair's high word and the tag are uniform across a wave.  Which words are
// uniform (U) or per-lane (L) entering each round:
//   round 0 (U, L, U, U), round 1 (L, U, U, U), round 2 (U, U, L, L),
//   round 3 (L, L, L, U), round 4 on (L, L, L, L).
// Rounds 0 and 1 keep plain xors (the compiler runs the uniform ones on the
// scalar unit); from round 2 on each output word is one bitop3.  In round 2
// both output xors, and in round 3 the z word's, combine one per-lane value
// with two uniform ones, so their bitop3 copies one scalar operand into a
// vector register first.  SPLIT xors the uniform pair on the scalar unit
// and then the per-lane value with one v_xor instead: three vector
// instructions fewer per call, but more scalar work and scalar-to-vector
// dependencies.  Same bits either way.  Measured (same box, alternated)
// with SPLIT in every call: the FP32 Euler kernels 1.2-1.8 % faster (Bates
// surface 5.16 -> 5.07-5.10 ms, uneven batch 20.3 -> 20.0 ms), the FP32 QE
// kernels 1 % slower, the FP64 QE kernels 0.5 % slower, FP64 Euler
// unchanged, and the latency-bound Heston Euler call (100k paths, one wave
// per SIMD, where every dependency shows) 6 % slower; so the normal draws of
// the step loops (philox_steps) split in the FP32 Bates Euler loop only, and
// the rare jump calls always.
template <bool SPLIT = true>
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)kPhiloxM0 * c.x;
        const uint64_t p1 = (uint64_t)kPhiloxM1 * c.z;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        if (r < 2)
            c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
        else if (SPLIT && r == 2)
            c = make_uint4(hi1 ^ (c.y ^ k0), lo1, c.w ^ (hi0 ^ k1), lo0);
        else if (SPLIT && r == 3)
            c = make_uint4(xor3(hi1, c.y, k0), lo1, hi0 ^ (c.w ^ k1), lo0);
        else
            c = make_uint4(xor3(hi1, c.y, k0), lo1, xor3(hi0, c.w, k1), lo0);
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
__device__ __forceinline__ float hw_rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float hw_sin_rev(float x) { return __builtin_amdgcn_sinf(x); }  // sin(2 pi x)
__device__ __forceinline__ float hw_cos_rev(float x) { return __builtin_amdgcn_cosf(x); }  // cos(2 pi x)
__device__ __forceinline__ float clamp_pos(float x) {                                     // max(x, 0), one v_med3
    return __builtin_amdgcn_fmed3f(x, 0.0f, 3.402823466e38f);
}
// A uniform constant kept in a vector register for a whole loop.  A VOP3
// instruction reads at most one scalar operand, so an FMA of two uniform
// constants and a per-lane value would otherwise copy one of the constants
// into a vector register at every step.
__device__ __forceinline__ float in_vgpr(float x) {
    asm("" : "+v"(x));
    return x;
}
__device__ __forceinline__ double in_vgpr(double x) {
    asm("" : "+v"(x));
    return x;
}

constexpr double kLog2e = 1.4426950408889634;
constexpr double kLn2 = 0.6931471805599453;

// --------------------------------------------------------- normal draws
// Box-Muller on two 32-bit words: u1 = (w0 + 1/2) 2^-32 in (0,1), u2 =
// (w1 >> 9) 2^-23 in [0,1) (a 23-bit angle: the FP32 cosine cannot use more,
// and the angle word's 9 low bits serve as the Poisson gate of the step);
// za = r cos(2 pi u2), zb = r sin(2 pi u2), r = sqrt(-2 ln u1).  Both
// precisions consume the same words, so FP32 and FP64 runs see the same
// normals up to rounding.
__device__ __forceinline__ void box_muller_f64(uint32_t w0, uint32_t w1, double& za, double& zb) {
    const double u1 = ((double)w0 + 0.5) * 0x1p-32;
    const double r = sqrt_f64(-2.0 * log_f64(u1));
    double s, c;
    sincos_rev23(w1 >> kGateBits, s, c);
    za = r * c;
    zb = r * s;
}

// 1 + u2 as an FP32 value in [1, 2), exactly: the angle bits w1 >> 9 become
// the mantissa below the exponent of 1.0 in one v_alignbit_b32.  The
// hardware cosine and sine take revolutions and have period 1, so they take
// 1 + u2 for u2.
__device__ __forceinline__ float angle_rev_f32(uint32_t w1) {
    return __uint_as_float(__builtin_amdgcn_alignbit(0x7Fu, w1, kGateBits));
}

__device__ __forceinline__ void box_muller_f32(uint32_t w0, uint32_t w1, float& za, float& zb) {
    const float u1 = fmaf((float)w0, 0x1p-32f, 0x1p-33f);
    const float a = angle_rev_f32(w1);
    const float r = hw_sqrt(-1.38629436111989061f * hw_log2(u1));  // -2 ln2 log2(u1)
    za = r * hw_cos_rev(a);
    zb = r * hw_sin_rev(a);
}

// Exact Poisson count by inverse CDF of (w + 1/2) 2^-32.  The common n = 0
// case is decided by one integer compare against a host-computed threshold
// (w < thr0 <=> u < e^-mu); this tail walks the CDF in FP64 and returns >= 1.
__device__ __forceinline__ uint32_t poisson_tail(uint32_t w, double mu, double e0) {
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

// Philox substream of one pair: key = seed, counter = (call, pair lo, pair hi, tag).
struct PairStream {
    uint32_t k0, k1, plo, phi;
    template <bool SPLIT = true>
    __device__ __forceinline__ uint4 call(uint32_t c0, uint32_t tag) const {
        return philox4x32_10<SPLIT>(make_uint4(c0, plo, phi, tag), k0, k1);
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

// Jump of step s given the 9 high bits h of its Poisson uniform word, for
// h >= thr0 >> 23 (the caller's test: below it the count is 0 whatever the
// low bits): jump_sum = mu_j n + sigma_j sqrt(n) z (mc.rs:139-140).  The
// jump call of step s supplies the low 23 bits (word 2), completing the
// 32-bit word W of the exact inverse-CDF count, and the jump normal (the
// Box-Muller cosine of words 0, 1) when the count is non-zero.  FP64 mode
// returns the jump in ln units, FP32 mode in log2 units.
template <bool FP32>
__device__ __forceinline__ auto jump_of_step(const McDesc& d, const PairStream& st, uint32_t s, uint32_t h,
                                             uint32_t& n) {
    using R = typename std::conditional<FP32, float, double>::type;
    R j = 0;
    n = 0;
    const uint4 w = st.call(s, kTagJump);
    const uint32_t W = (h << (32u - kGateBits)) | (w.z >> kGateBits);
    if (W >= d.pois_thr0) {
        n = W < d.pois_thr1 ? 1u : poisson_tail(W, d.pois_mu, d.pois_e0);
        if constexpr (FP32) {
            float z, unused;
            box_muller_f32(w.x, w.y, z, unused);
            j = fmaf(d.f_sigma_j2 * hw_sqrt((float)n), z, d.f_mu_j2 * (float)n);
        } else {
            double z, unused;
            box_muller_f64(w.x, w.y, z, unused);
            j = d.mu_j * (double)n + d.sigma_j * sqrt_f64((double)n) * z;
        }
    }
    return j;
}

struct NoBlockEnd {
    __device__ void operator()(uint32_t, PairStream&) const {}
};

// Drives the Philox-mode draws of one pair through all steps, calling
// step(za, zb, jump, n, s) once per step s: one Philox call per two steps.
// Bates: the 9 low bits of a step's angle word are the high bits of its
// Poisson uniform; below thr0 >> 23 (probability e^{-lambda dt} up to 2^-9)
// the count is 0, and only otherwise does the step draw its jump call.
//
// DEFER (European payoffs, whose log-price needs only the sum of the
// jumps): step() gets jump 0, and a step whose gate bits allow a count
// only records (step, gate bits) in one of three per-lane slots; the jumps
// of the slots are drawn after the last step and added to *jsum.  A wave then
// runs the rare jump code about three times per path instead of at every
// step where one of its lanes may jump.  A lane's fourth and later events
// are drawn at once.  The draws are the same words either way.  (The slots
// measured faster in registers than as one LDS store per event.)
//
// block_end(s, st) is called after step s for every s = 8k + 7 of the steps
// that complete a block of eight (FP32 mode, in which the step sums its
// increments in FP32 over such blocks, and FP64 mode when a block_end is
// given); the loops run whole blocks, so the call sits at the end of a loop
// body rather than behind a per-step branch, whose two sides would need
// register copies.  block_end may hand the lane another pair's state and
// stream (st, the pair index words; see the QE regrouping).
//
// MIX (FP32 Euler): step() gets the scaled, correlated normals of the Euler
// step, sa = a2 za and sb = b2 (rho za + sqrt(1 - rho^2) zb), instead of
// (za, zb).  With za = r cos t and zb = r sin t, rho = cos phi and
// sqrt(1 - rho^2) = sin phi, the mixed normal is r cos(t - phi): one cosine
// replaces the sine and the mixing products, and a2, b2 fold into r.
template <bool FP32, bool JUMPS, bool DEFER = false, bool MIX = false, class Step, class BlockEnd = NoBlockEnd>
__device__ __forceinline__ void philox_steps(const McDesc& d, const PairStream& st0, Step&& step,
                                             typename std::conditional<FP32, float, double>::type* jsum = nullptr,
                                             BlockEnd block_end = BlockEnd{}) {
    using R = typename std::conditional<FP32, float, double>::type;
    static_assert(!DEFER || JUMPS, "deferred jumps need the jump model");
    static_assert(!MIX || FP32, "mixed normals are an FP32 mode");
    constexpr bool kBlocks = FP32 || !std::is_same<BlockEnd, NoBlockEnd>::value;
    PairStream st = st0;
    // only the FP32 Bates Euler loop's normal draws split the xors (see
    // philox4x32_10)
    constexpr bool kSplit = MIX && JUMPS;
    // DEFER: the lane's recorded events, oldest first, as (step, gate bits);
    // step ~0u marks an empty slot
    uint32_t es0 = ~0u, ew0 = 0u, es1 = ~0u, ew1 = 0u, es2 = ~0u, ew2 = 0u;
    R jd = 0;
    const float phi = MIX ? in_vgpr(d.eu_phi) : 0.0f;
    const uint32_t gate = d.pois_gate;
    auto one = [&](uint32_t wa, uint32_t wb, uint32_t s) {
        R za, zb;
        if constexpr (MIX) {
            const float u1 = fmaf((float)wa, 0x1p-32f, 0x1p-33f);
            const float a = angle_rev_f32(wb);
            const float ra = hw_sqrt(d.eu_r2 * hw_log2(u1));  // a2 sqrt(-2 ln u1)
            za = ra * hw_cos_rev(a);
            zb = (ra * d.eu_kb) * hw_cos_rev(a - phi);
        } else if constexpr (FP32) {
            box_muller_f32(wa, wb, za, zb);
        } else {
            box_muller_f64(wa, wb, za, zb);
        }
        uint32_t n = 0;
        R j = 0;
        if constexpr (JUMPS) {
            const uint32_t h = wb & kGateMask;
            if (h >= gate) {  // rare: P ~ 1 - e^{-lambda dt} + 2^-9
                if constexpr (DEFER) {
                    if (es0 == ~0u) {
                        es0 = s;
                        ew0 = h;
                    } else if (es1 == ~0u) {
                        es1 = s;
                        ew1 = h;
                    } else if (es2 == ~0u) {
                        es2 = s;
                        ew2 = h;
                    } else {
                        uint32_t nn;
                        jd += jump_of_step<FP32>(d, st, s, h, nn);
                    }
                } else {
                    j = jump_of_step<FP32>(d, st, s, h, n);
                }
            }
        }
        step(za, zb, j, n, s);
    };
    const uint32_t N = d.n_steps;
    uint32_t i = 0;
    if constexpr (kBlocks) {
        // blocks of eight steps; FP32 Euler unrolls a block (the scheduler
        // overlaps consecutive steps), the QE step, three times longer,
        // measured slower unrolled and keeps a two-step body
        for (; i + 8 <= N; i += 8) {
            if constexpr (MIX) {
#pragma unroll
                for (uint32_t k = 0; k < 8; k += 2) {
                    const uint4 w = st.template call<kSplit>((i + k) >> 1, kTagNormal);
                    one(w.x, w.y, i + k);
                    one(w.z, w.w, i + k + 1);
                }
            } else {
#pragma nounroll
                for (uint32_t k = 0; k < 8; k += 2) {
                    const uint4 w = st.template call<kSplit>((i + k) >> 1, kTagNormal);
                    one(w.x, w.y, i + k);
                    one(w.z, w.w, i + k + 1);
                }
            }
            block_end(i + 7, st);
        }
    }
    for (; i + 2 <= N; i += 2) {
        const uint4 w = st.template call<kSplit>(i >> 1, kTagNormal);
        one(w.x, w.y, i);
        one(w.z, w.w, i + 1);
    }
    if (i < N) {
        const uint4 w = st.template call<kSplit>(i >> 1, kTagNormal);
        one(w.x, w.y, i);
    }
    if constexpr (DEFER) {
        uint32_t nn;
        if (es0 != ~0u) jd += jump_of_step<FP32>(d, st, es0, ew0, nn);
        if (es1 != ~0u) jd += jump_of_step<FP32>(d, st, es1, ew1, nn);
        if (es2 != ~0u) jd += jump_of_step<FP32>(d, st, es2, ew2, nn);
        *jsum += jd;
    }
}

// ------------------------------------ QE: regrouping pairs by variance
// QE's exponential branch serves the pairs whose variance is close to 0 (psi
// = s^2/m^2 > 1.5), a few in a hundred at the parameters that matter; but
// one such lane makes its whole wave run the branch, which happens in ~90 %
// of wave-steps when pairs stay on their lanes: a fifth of the FP32 QE
// kernel's time, two fifths of the FP64 one's.  Low variance persists over
// steps, so every 8 steps (at the block ends of the step driver) the tile
// block regroups its pairs: the lanes whose variance is below the host's
// threshold (u = psi/2 > 0.1 there; mc_desc.h qe_vlow) move to the front
// rows in wave and lane order, the others follow, the order rotated by 64
// rows per block and per regrouping so that no SIMD always holds the low
// wave.  A lane hands over its whole pair (state words and Philox pair
// index) through LDS; after the last step every pair returns to its home
// lane.  A pair's steps are the same instructions on the same values
// wherever it runs, so results do not change; waves with no low lane skip
// the branch (57 % of wave-steps run it instead of 90 %, in a CPU model of
// the variance paths at v0 = theta = 0.04, kappa = 1.5, sigma = 0.5, 365
// steps; measured: FP32 QE pair 1.08 -> 1.04 ms, FP64 5.32 -> 5.07 ms, of
// which the regrouping itself costs ~3 %).  Every thread of the block must
// step (two __syncthreads per regrouping).
constexpr int kGroupQuads = 4;  // most 16-byte quads of state a lane hands over (FP64 Asian / up-and-out)
struct GroupLds {
    uint4 rows[kGroupQuads][kBlock];
    uint32_t cnt[kBlock / 64];
};

// Row this lane's pair moves to: low lanes first (wave order, then lane
// order), then the rest, rotated by 64 rot rows.  One __syncthreads.
__device__ __forceinline__ uint32_t group_dest(GroupLds* g, bool low, uint32_t rot) {
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint64_t m = __ballot(low);
    const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
    if (lane == 0) g->cnt[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    const uint32_t c0 = g->cnt[0], c1 = g->cnt[1], c2 = g->cnt[2], c3 = g->cnt[3];
    const uint32_t pre = (wave > 0 ? c0 : 0u) + (wave > 1 ? c1 : 0u) + (wave > 2 ? c2 : 0u);
    const uint32_t dst = low ? pre + rank : (c0 + c1 + c2 + c3) + (64u * wave - pre) + (lane - rank);
    return (dst + 64u * rot) & (uint32_t)(kBlock - 1);
}

// Writes rec (W words, in 16-byte quads) to row dst and takes this thread's
// row.  One __syncthreads; the rows are free again once every thread has
// passed the next barrier.
template <int W>
__device__ __forceinline__ void group_move(GroupLds* g, uint32_t dst, uint32_t (&rec)[W]) {
    static_assert(W % 4 == 0 && W / 4 <= kGroupQuads, "state quads");
#pragma unroll
    for (int q = 0; q < W / 4; ++q) g->rows[q][dst] = make_uint4(rec[4 * q], rec[4 * q + 1], rec[4 * q + 2], rec[4 * q + 3]);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < W / 4; ++q) {
        const uint4 v = g->rows[q][threadIdx.x];
        rec[4 * q] = v.x;
        rec[4 * q + 1] = v.y;
        rec[4 * q + 2] = v.z;
        rec[4 * q + 3] = v.w;
    }
}

__device__ __forceinline__ void put_f64(uint32_t* r, double x) {
    const uint64_t b = (uint64_t)__double_as_longlong(x);
    r[0] = (uint32_t)b;
    r[1] = (uint32_t)(b >> 32);
}
__device__ __forceinline__ double get_f64(const uint32_t* r) {
    return __longlong_as_double((long long)(((uint64_t)r[1] << 32) | r[0]));
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
// transcendentals in FP32 on the hardware units.  Loop invariants are folded
// into log2 units on the host.
//
// Euler for one sibling in FP32: sa = a2*z1 and sb = b2*z2 are per-pair
// products (the antithetic sibling passes them negated).  v >= 0 on entry:
// every step ends with max(., 0), and a negative v0 enters as 0 (see
// sim_pair_f32).  So v+ = v and v + b0 + b1 v+ = c1 v + b0 with c1 = 1 + b1
// (c1 passed by the caller, held in a vector register).  Returns
// acc + a1 v + sqrt(v) sa: the caller passes a0 (+ the jump) for the
// increment itself, or a running partial sum to which the constant a0 is
// added once per path.
__device__ __forceinline__ float euler_f32(const McDesc& d, float& v, float sa, float sb, float acc, float c1,
                                           float b0) {
    const float sv = hw_sqrt(v);
    const float r = fmaf(sv, sa, fmaf(d.eu_a1, v, acc));
    v = fmaxf(fmaf(sv, sb, fmaf(c1, v, b0)), 0.0f);
    return r;
}

// ncdf(-|z|) = erfc(x)/2, x = |z|/sqrt2, from erfc(x) = t exp(P(t) - x^2),
// t = 1/(1 + x/2), which has fractional error below 1.2e-7 for every x >= 0:
// the Chebyshev fit of Press, Teukolsky, Vetterling and Flannery, Numerical
// Recipes (2nd ed.), section 6.2, erfcc.  Evaluated as t 2^(Q(t) - xs^2) with
// Q = log2e P - 1 (the halving) and xs = x sqrt(log2e), so that no separate
// scaling multiplies are needed.
__device__ __forceinline__ float ncdf_neg_abs_f32(float z) {
    constexpr double L = 1.4426950408889634;                               // log2 e
    const float xs = fabsf(z) * 0.8493218002880191f;                       // |z| sqrt(log2e / 2)
    const float t = hw_rcp(fmaf(xs, 0.41627730557884884f, 1.0f));          // 1/(1 + x/2)
    float p = (float)(0.17087277 * L);
    p = fmaf(p, t, (float)(-0.82215223 * L));
    p = fmaf(p, t, (float)(1.48851587 * L));
    p = fmaf(p, t, (float)(-1.13520398 * L));
    p = fmaf(p, t, (float)(0.27886807 * L));
    p = fmaf(p, t, (float)(-0.18628806 * L));
    p = fmaf(p, t, (float)(0.09678418 * L));
    p = fmaf(p, t, (float)(0.37409196 * L));
    p = fmaf(p, t, (float)(1.00002368 * L));
    p = fmaf(p, t, (float)(-1.26551223 * L - 1.0));
    return t * hw_exp2(fmaf(-xs, xs, p));
}

// QE variance step in FP32 (Andersen 2008, psi_c = 1.5) for one sibling, in
// terms of the conditional mean m, s2h = s2/2 (the host halves its
// coefficients) and u = psi/2 = s2h/m^2; psi > 1.5 <=> u > 0.75 <=> s2h >
// 0.75 m^2.
// * quadratic: with w = sqrt(1 - u), 2/psi = 1/u and
//   b^2 = 2/psi - 1 + sqrt(2/psi (2/psi - 1)) = w (1 + w)/u,
//   a = m/(1 + b^2) = m u/(1 + w), so a b^2 = m w and
//   v' = a (b + z)^2 = m w (1 + z/b)^2.  With D = m^2 - s2h = (m w)^2
//   (m >= 0), m w = sqrt(D) and w (1 + w)/u = m w (m + m w)/s2h, so
//   1/b = sqrt(s2h/(m w (m + m w))) = s2h rsqrt(s2h (m w m + D)):
//   two transcendentals, no division and no u, and no cancellation for u in
//   [0, 0.75] (D >= m^2/4).  FLT_MIN added to the rsqrt's argument keeps
//   s2h = 0 (sigma = 0, or theta = v = 0) at 1/b = 0, v' = m, instead of
//   0 * inf; the reference's psi floor (psi >= 1e-12) moves v' by at most
//   ~1e-6 relative there.  Below m ~ 1e-7 the argument underflows and 1/b
//   loses precision, where v' itself is ~1e-7 and the price does not see it;
// * exponential: 1 - p = 2/(psi + 1) = 1/(u + 1/2) and 1/beta = m (u + 1/2),
//   and with phic = 1 - ncdf(z): v' = 0 when u_z <= p (phic (u + 1/2) >= 1),
//   else ln((1 - p)/phic)/beta = m (u + 1/2) max(-ln((u + 1/2) phic), 0),
//   which avoids the cancellation of 1 - ncdf(z).  Only this branch needs
//   u; its lanes have s2h > 0.75 m^2 >= 0, so u > 0.75 (the psi floor
//   cannot act) or u = s2h / 0 = inf, capped at 1e30 so that a vanishing
//   m^2 leaves every product finite.
// qe_sib_f32 returns m, s2h and the quadratic branch's v'; on exponential
// lanes that v' is meaningless (sqrt of D < m^2/4) and the caller replaces it.
struct QeSib {
    float m, s2h, vq;
    bool ex;
};

__device__ __forceinline__ QeSib qe_sib_f32(const McDesc& d, float v, float z, float m0, float s0) {
    QeSib o;
    o.m = fmaf(d.qe_m1, v, m0);
    o.s2h = fmaf(d.qe_s1, v, s0);
    const float mm = o.m * o.m;
    o.ex = o.s2h > 0.75f * mm;
    const float D = mm - o.s2h;
    const float mw = hw_sqrt(D);  // m w
    const float ib = o.s2h * hw_rsq(fmaf(o.s2h, fmaf(mw, o.m, D), 0x1p-126f));  // 1/b
    const float f = fmaf(z, ib, 1.0f);
    o.vq = mw * (f * f);
    return o;
}

// The exponential branch's v' for tail probability phic = 1 - ncdf(z).
__device__ __forceinline__ float qe_exp_f32(const QeSib& s, float phic) {
    const float u = fminf(s.s2h * hw_rcp(s.m * s.m), 1e30f);
    const float uh = u + 0.5f;
    return (s.m * uh) * fmaxf(hw_log2(uh * phic) * -0.69314718055994531f, 0.0f);
}

// One QE step of both siblings of a pair: sibling a draws (za, zb), sibling b
// (-za, -zb), when the pair has one (ANTI, fixed per compiled loop: a
// runtime flag costs register copies and selects at every step).  |za| is
// common to both, so the exponential branch's erfc is evaluated once per
// pair, and only when a lane of the wave needs it; its results replace the
// quadratic branch's v' inside that branch, so that nothing has to be
// initialised on the way round it.
// The variances are >= 0 on entry (QE's v' is never negative, and v0 is
// clamped by the caller).  The price variance term K3 v + K4 v' has K3 = K4
// (gamma1 = gamma2 = 1/2), so sqrt(max(K3 v + K4 v', 0)) = sqrt(max(K3, 0))
// sqrt(v + v'), and the constant folds into zb.  m0, s0 and q0 are uniform
// constants held in vector registers by the caller.  Returns the log2-price
// increments.
template <bool JUMPS, bool ANTI>
__device__ __forceinline__ void qe_pair_f32(const McDesc& d, float& va, float& vb, float za, float zb, float jl,
                                            float m0, float s0, float q0, float& dya, float& dyb) {
    const QeSib a = qe_sib_f32(d, va, za, m0, s0);
    QeSib b{};
    if (ANTI) b = qe_sib_f32(d, vb, -za, m0, s0);
    float na = a.vq, nb = b.vq;
    if (a.ex || (ANTI && b.ex)) {
        // z = za for sibling a, z = -za for sibling b: 1 - ncdf(z) is e =
        // ncdf(-|za|) for the sibling whose z >= 0 and 1 - e for the other
        const float e = ncdf_neg_abs_f32(za), ome = 1.0f - e;
        const bool pos = za >= 0.0f;
        if (a.ex) na = qe_exp_f32(a, pos ? e : ome);
        if (ANTI && b.ex) nb = qe_exp_f32(b, pos ? ome : e);
    }
    const float zs = d.qe_sq3 * zb;
    auto incr = [&](float vp, float vn, float z) {
        float dy = fmaf(hw_sqrt(vp + vn), z, fmaf(d.qe_q2, vn, fmaf(d.qe_q1, vp, q0)));
        if (JUMPS) dy += jl;
        return dy;
    };
    dya = incr(va, na, zs);
    va = na;
    if (ANTI) {
        dyb = incr(vb, nb, -zs);
        vb = nb;
    } else {
        dyb = 0.0f;
    }
}

// ------------------------------------------------------------ path state
struct PathState {
    double ln_s;   // ln S_T
    double v;      // v_T
    double sum;    // running_sum (Asian only, else NaN)
    double max;    // running_max (up-and-out only, else NaN)
    double x;      // payoff underlying: S_T, or running_sum / n_steps for Asian
    int ko;        // knocked out (up-and-out)
    int near;      // FP32 path: running maximum within kBarrierMargin of the barrier
};

// FP32-path up-and-out: a path whose running maximum of log2(S/S0) lies
// within this distance of log2(B/S0) is decided by its FP64 path.  The FP32
// and FP64 maxima of one path differ by at most ~3e-6 in 99.9 % of paths
// (QE and Euler, 252-500 steps, measured on identical draws).
constexpr double kBarrierMargin = 0x1p-16;

__device__ __forceinline__ double payoff_value(const McDesc& d, const PathState& e, double strike) {
    // disc * Payoff::eval (mc.rs:75-95, 275-277)
    const double intr = d.opt_put ? fmax(strike - e.x, 0.0) : fmax(e.x - strike, 0.0);
    return d.disc * (e.ko ? d.rebate : intr);
}

// Payoff::eval before discounting, for the strike loop: sgn = +1 (call) or -1
// (put) and nk = -sgn K, so fma(sgn, x, nk) is x - K or K - x with the one
// rounding of the subtraction, without evaluating both and selecting.
__device__ __forceinline__ double payoff_undiscounted(const McDesc& d, const PathState& e, double sgn, double nk) {
    const double intr = fmax(fma(sgn, e.x, nk), 0.0);
    return e.ko ? d.rebate : intr;
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

// ------------------------------------------- FP64 path, Philox mode
// The Philox-mode FP64 kernels need not repeat the reference's evaluation
// order (their draws are not the reference's; replay mode keeps that order),
// so they take the FP32 path's algebra in FP64: loop invariants hoisted on
// the host (McDesc ed_* / qd_*), no division by psi, one erfc per antithetic
// pair, reciprocals from rcp_f64, and the fast log / exp / sincos of
// fp64_math.h.  Results agree with the reference-order arithmetic to a few
// ulp per step.

// Full-truncation Euler for one sibling: v >= 0 on entry (a negative v0
// enters as 0 with b0 + v0 for its first step, as in the FP32 path).  sa =
// sqrt(dt) z1 and sb = sigma sqrt(dt) z2 are per-pair (negated for sibling b).
// a0 = d.ed_a0, held in a vector register by the caller.
__device__ __forceinline__ double euler_f64_fast(const McDesc& d, double& v, double sa, double sb, double b0,
                                                 double a0) {
    const double sv = sqrt_f64(v);
    const double inc = fma(sv, sa, fma(d.ed_a1, v, a0));
    v = fmax(fma(sv, sb, fma(d.ed_c1, v, b0)), 0.0);
    return inc;
}

// The QE variance step of one sibling in FP64, as qe_sib_f32: m, s2h and the
// quadratic branch's v' = m w (1 + z/b)^2 with m w = sqrt(D), D = m^2 - s2h,
// and 1/b = s2h rsqrt(s2h (m w m + D)) (DBL_MIN added against 0 * inf at
// s2h = 0); the exponential branch forms u = psi/2 (> 0.75 there, capped at
// 1e30).
struct QeSib64 {
    double m, s2h, vq;
    bool ex;
};

// m0 = d.qd_m0 and s0 = d.qd_s0.
__device__ __forceinline__ QeSib64 qe_sib_f64(const McDesc& d, double v, double z, double m0, double s0) {
    QeSib64 o;
    o.m = fma(d.ekt, v, m0);
    o.s2h = fma(d.qd_s1, v, s0);
    const double mm = o.m * o.m;
    o.ex = o.s2h > 0.75 * mm;
    const double D = mm - o.s2h;
    const double mw = sqrt_f64(D);
    const double ib = o.s2h * rsqrt_f64(fma(o.s2h, fma(mw, o.m, D), 0x1p-1022));
    const double f = fma(z, ib, 1.0);
    o.vq = mw * (f * f);
    return o;
}

// The exponential branch's v' for tail probability phic = 1 - ncdf(z).
__device__ __forceinline__ double qe_exp_f64(const QeSib64& s, double phic) {
    const double u = fmin(s.s2h * rcp_f64(s.m * s.m), 1e30);
    const double uh = u + 0.5;
    return (s.m * uh) * fmax(-log_f64(uh * phic), 0.0);
}

// One QE step of both siblings (mc.rs:256-267, 288-309): returns the ln-price
// increments (without the jump).  |za| is common to the siblings, so the
// exponential branch's erfc is evaluated once per pair, and only when a lane
// of the wave needs it.  K3 = K4, so sqrt(max(K3 v + K4 v', 0)) =
// sqrt(max(K3, 0)) sqrt(v + v').
// m0, s0 and a0 are d.qd_m0, d.qd_s0 and d.qe_a0.
__device__ __forceinline__ void qe_pair_f64_fast(const McDesc& d, bool anti, double& va, double& vb, double za,
                                                 double zb, double m0, double s0, double a0, double& dya,
                                                 double& dyb) {
    const QeSib64 a = qe_sib_f64(d, va, za, m0, s0);
    QeSib64 b{};
    if (anti) b = qe_sib_f64(d, vb, -za, m0, s0);
    double na = a.vq, nb = b.vq;
    if (a.ex || (anti && b.ex)) {
        // e = ncdf(-|za|) (math.rs:16-19); 1 - ncdf(z) is e for the sibling
        // whose z >= 0 and 1 - e for the other
        const double e = ncdf_neg_abs_f64(za), ome = 1.0 - e;
        const bool pos = za >= 0.0;
        if (a.ex) na = qe_exp_f64(a, pos ? e : ome);
        if (anti && b.ex) nb = qe_exp_f64(b, pos ? ome : e);
    }
    const double zs = d.qd_sq3 * zb;
    dya = fma(sqrt_f64(va + na), zs, fma(d.k2, na, fma(d.k1, va, a0)));
    va = na;
    if (anti) {
        dyb = fma(sqrt_f64(vb + nb), -zs, fma(d.k2, nb, fma(d.k1, vb, a0)));
        vb = nb;
    } else {
        dyb = 0.0;
    }
}

// Pair simulation, FP64, Philox mode.  GROUP: the block regroups its pairs
// by variance every 8 steps (as the FP32 QE loop).
template <int SCHEME, bool JUMPS, int PAYOFF, bool GROUP>
__device__ __forceinline__ void sim_pair_f64_fast(const McDesc& d, const PhiloxSource& src, bool anti,
                                                  PathState& ea, PathState& eb, void* lds) {
    double la = d.ln_spot, lb = d.ln_spot;
    double va = fmax(d.v0, 0.0), vb = va;
    double suma = 0.0, sumb = 0.0, jsum = 0.0;
    double mxa = -__builtin_huge_val(), mxb = -__builtin_huge_val();
    const uint32_t N = d.n_steps;
    double b0s = d.v0 < 0.0 ? d.v0 + d.ed_b0 : d.ed_b0;
    constexpr bool kDefer = JUMPS && PAYOFF == kPayEuropean;
    // Euler: the uniform addend of an FMA whose multiplier is uniform too
    // (one scalar operand per FMA) held in a vector register for the whole
    // loop (Bates surface FP64 15.52 -> 15.30 ms).  The QE loop passes its
    // addends as they are: held in vector registers they pushed it from 137
    // to 168 VGPRs with spills, for no gain.
    const double ea0 = SCHEME == kSchemeEuler ? in_vgpr(d.ed_a0) : 0.0;

    auto step = [&](double za, double zb, double jump, uint32_t, uint32_t) {
        double dya, dyb;
        if (SCHEME == kSchemeEuler) {
            const double sa = d.sqdt * za;
            const double sb = d.ed_sb * fma(d.rho, za, d.sqrt_1m_rho2 * zb);
            dya = euler_f64_fast(d, va, sa, sb, b0s, ea0);
            dyb = euler_f64_fast(d, vb, -sa, -sb, b0s, ea0);  // sibling b unused without antithetic pairs
            b0s = d.ed_b0;
        } else {
            qe_pair_f64_fast(d, anti, va, vb, za, zb, d.qd_m0, d.qd_s0, d.qe_a0, dya, dyb);
        }
        if (JUMPS && !kDefer) {
            dya += jump;
            dyb += jump;
        }
        la += dya;
        lb += dyb;
        if (PAYOFF == kPayAsian) {
            suma += exp_f64(la);
            sumb += exp_f64(lb);
        }
        if (PAYOFF == kPayUpAndOut) {
            mxa = fmax(mxa, la);
            mxb = fmax(mxb, lb);
        }
    };
    if constexpr (GROUP) {
        static_assert(SCHEME == kSchemeQE && !kDefer, "regrouping: QE without deferred jumps");
        GroupLds* g = static_cast<GroupLds*>(lds);
        uint32_t home = threadIdx.x;
        constexpr int W = 12 + (PAYOFF == kPayAsian || PAYOFF == kPayUpAndOut ? 4 : 0);
        auto move = [&](uint32_t dst, PairStream& st) {
            uint32_t r[W] = {};
            put_f64(r + 0, la);
            put_f64(r + 2, lb);
            put_f64(r + 4, va);
            put_f64(r + 6, vb);
            r[8] = st.plo;
            r[9] = home;
            if constexpr (PAYOFF == kPayAsian) {
                put_f64(r + 10, suma);
                put_f64(r + 12, sumb);
            }
            if constexpr (PAYOFF == kPayUpAndOut) {
                put_f64(r + 10, mxa);
                put_f64(r + 12, mxb);
            }
            group_move(g, dst, r);
            la = get_f64(r + 0);
            lb = get_f64(r + 2);
            va = get_f64(r + 4);
            vb = get_f64(r + 6);
            st.plo = r[8];
            home = r[9];
            if constexpr (PAYOFF == kPayAsian) {
                suma = get_f64(r + 10);
                sumb = get_f64(r + 12);
            }
            if constexpr (PAYOFF == kPayUpAndOut) {
                mxa = get_f64(r + 10);
                mxb = get_f64(r + 12);
            }
        };
        auto block_end = [&](uint32_t s, PairStream& st) {
            const double vk = anti ? fmin(va, vb) : va;
            move(group_dest(g, vk < (double)d.qe_vlow, (blockIdx.x + (s >> 3)) & 3u), st);
        };
        philox_steps<false, JUMPS, kDefer>(d, src.st, step, &jsum, block_end);
        PairStream st = src.st;  // every pair back to its home lane
        __syncthreads();
        move(home, st);
    } else {
        philox_steps<false, JUMPS, kDefer>(d, src.st, step, &jsum);
    }
    if (kDefer) {
        la += jsum;
        lb += jsum;
    }

    const double nan = __builtin_nan("");
    auto finish = [&](PathState& e, double ln_s, double v, double sum, double mx) {
        e.ln_s = ln_s;
        e.v = v;
        e.sum = PAYOFF == kPayAsian ? sum : nan;
        e.max = nan;
        e.ko = 0;
        e.near = 0;
        if (PAYOFF == kPayAsian) {
            e.x = sum / (double)N;
        } else {
            e.x = exp(ln_s);
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

// --------------------------------------------- pair simulation, FP64 path
// Replay mode runs the reference's expressions in the reference's order
// (mc.rs:243-277); Philox mode the hoisted form above.
// GROUP: as sim_pair_f64_fast (lds: the block's LDS).
template <int SCHEME, bool JUMPS, int PAYOFF, bool GROUP = false, class Src>
__device__ __forceinline__ void sim_pair_f64(const McDesc& d, const Src& src, bool anti, PathState& ea,
                                             PathState& eb, void* lds = nullptr) {
    if constexpr (Src::kPhilox) {
        sim_pair_f64_fast<SCHEME, JUMPS, PAYOFF, GROUP>(d, src, anti, ea, eb, lds);
        return;
    }
    double la = d.ln_spot, lb = d.ln_spot, va = d.v0, vb = d.v0;
    double suma = 0.0, sumb = 0.0;
    double mxa = -__builtin_huge_val(), mxb = -__builtin_huge_val();
    const uint32_t N = d.n_steps;

    auto step = [&](double za, double zb, double jump, uint32_t, uint32_t) {
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
            // both unconditionally (mxb is unused without the sibling): a
            // conditional second update is merged into one store through a
            // selected address, which keeps the pair in scratch memory
            mxa = fmax(mxa, la);
            mxb = fmax(mxb, lb);
        }
    };

    if constexpr (!Src::kPhilox) {
        for (uint32_t i = 0; i < N; ++i) {
            double za, zb, j;
            src.get(i, za, zb, j);
            step(za, zb, j, 0u, i);
        }
    }

    const double nan = __builtin_nan("");
    auto finish = [&](PathState& e, double ln_s, double v, double sum, double mx) {
        e.ln_s = ln_s;
        e.v = v;
        e.sum = PAYOFF == kPayAsian ? sum : nan;
        e.max = nan;
        e.ko = 0;
        e.near = 0;
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
// ANTI: the pair has a sibling (QE's loop is compiled once per case; the
// Euler loop always computes the sibling, and `anti` decides whether its
// result is kept).  GROUP: the block regroups its pairs by variance every 8
// steps through lds (QE, Philox mode, no deferred jumps; every thread of
// the block steps).
template <int SCHEME, bool JUMPS, int PAYOFF, bool ANTI, bool GROUP, class Src>
__device__ __forceinline__ void sim_pair_f32_loop(const McDesc& d, const Src& src, bool anti, PathState& ea,
                                                  PathState& eb, void* lds) {
    double ya = 0.0, yb = 0.0;
    // QE only ever reads v+ = max(v, 0) and never produces a negative v; the
    // Euler step's handling of a negative v0 is below
    float va = clamp_pos(d.f_v0), vb = va;
    float suma = 0.0f, sumb = 0.0f;
    double mxa = -__builtin_huge_val(), mxb = -__builtin_huge_val();
    const uint32_t N = d.n_steps;

    // Increments are summed in FP32 over blocks of 8 steps (one per Poisson
    // call of the Bates draws), pa / pb, and folded into the FP64 totals
    // ya / yb once per block (block_end of the step driver).  European
    // payoffs need only y_T (Euler: the partial sums leave out the constant
    // a0, added as N a0 after the loop, and the jumps, which both siblings
    // share and which are summed once, in jsum).  Asian payoffs read y =
    // ya + pa every step through fa = (float)ya: S is exp2(fa + pa).
    // Barrier payoffs need the maximum of y over the steps: the block keeps
    // the maximum of its partial sums, ma / mb, and the fold takes ya +
    // (double)ma, the block's highest y, into the FP64 maxima mxa / mxb.
    float pa = 0.0f, pb = 0.0f, jsum = 0.0f;
    float fa = 0.0f, fb = 0.0f;
    float ma = -__builtin_huge_valf(), mb = -__builtin_huge_valf();
    auto fold = [&](uint32_t) {
        if (PAYOFF == kPayUpAndOut) {
            mxa = fmax(mxa, ya + (double)ma);
            mxb = fmax(mxb, yb + (double)mb);
            ma = mb = -__builtin_huge_valf();
        }
        ya += (double)pa;
        yb += (double)pb;
        pa = pb = 0.0f;
        if (PAYOFF == kPayAsian) {
            fa = (float)ya;
            fb = (float)yb;
        }
    };
    // A negative v0 has v+ = 0 in the first Euler step (mc.rs:245-254): no
    // diffusion and v' = max(v0 + b0, 0).  The general step computes exactly
    // that from v = 0 with b0 + v0 in place of b0, so the first step uses b0s
    // and every later one b0.
    float b0s = d.f_v0 < 0.0f ? d.f_v0 + d.eu_b0 : d.eu_b0;
    const float c1 = in_vgpr(d.eu_c1);
    const float q0 = in_vgpr(d.qe_q0), m0 = in_vgpr(d.qe_m0), s0 = in_vgpr(d.qe_s0);
    constexpr bool kMixed = Src::kPhilox && SCHEME == kSchemeEuler;
    auto step = [&](float za, float zb, float jl, uint32_t, uint32_t s) {
        if (SCHEME == kSchemeEuler) {
            // Philox mode passes sa, sb already mixed and scaled; sibling b is
            // computed whether or not it exists (its state is then unused),
            // which costs nothing more than selecting it out
            const float sa = kMixed ? za : d.eu_a2 * za;
            const float sb = kMixed ? zb : d.eu_b2 * fmaf(d.f_rho, za, d.f_rhoc * zb);
            if (PAYOFF == kPayEuropean) {
                if (JUMPS) jsum += jl;
                pa = euler_f32(d, va, sa, sb, pa, c1, b0s);
                pb = euler_f32(d, vb, -sa, -sb, pb, c1, b0s);
            } else {
                const float c = JUMPS ? d.eu_a0 + jl : d.eu_a0;
                pa += euler_f32(d, va, sa, sb, c, c1, b0s);
                pb += euler_f32(d, vb, -sa, -sb, c, c1, b0s);
            }
            b0s = d.eu_b0;
        } else {
            float dya, dyb;
            qe_pair_f32<JUMPS, ANTI>(d, va, vb, za, zb, jl, m0, s0, q0, dya, dyb);
            pa += dya;
            pb += dyb;  // dyb = 0 without the sibling
        }
        // sibling b's running state is updated unconditionally (unused without
        // the sibling): see sim_pair_f64
        if (PAYOFF == kPayAsian) {
            suma += hw_exp2(fa + pa);
            sumb += hw_exp2(fb + pb);
        }
        if (PAYOFF == kPayUpAndOut) {
            ma = fmaxf(ma, pa);
            mb = fmaxf(mb, pb);
        }
    };

    if constexpr (Src::kPhilox) {
        constexpr bool kDefer = JUMPS && PAYOFF == kPayEuropean;
        static_assert(!GROUP || (SCHEME == kSchemeQE && !kDefer), "regrouping: QE without deferred jumps");
        GroupLds* g = static_cast<GroupLds*>(lds);
        uint32_t home = threadIdx.x;
        // the words of a pair's state at a block end (pa = pb = 0, ma = mb =
        // -inf there; fa, fb follow from ya, yb)
        constexpr int W = 8 + (PAYOFF == kPayAsian || PAYOFF == kPayUpAndOut ? 4 : 0);
        auto move = [&](uint32_t dst, PairStream& st) {
            uint32_t r[W] = {};
            r[0] = __float_as_uint(va);
            r[1] = __float_as_uint(vb);
            put_f64(r + 2, ya);
            put_f64(r + 4, yb);
            r[6] = st.plo;
            r[7] = home;
            if constexpr (PAYOFF == kPayAsian) {
                r[8] = __float_as_uint(suma);
                r[9] = __float_as_uint(sumb);
            }
            if constexpr (PAYOFF == kPayUpAndOut) {
                put_f64(r + 8, mxa);
                put_f64(r + 10, mxb);
            }
            group_move(g, dst, r);
            va = __uint_as_float(r[0]);
            vb = __uint_as_float(r[1]);
            ya = get_f64(r + 2);
            yb = get_f64(r + 4);
            st.plo = r[6];
            home = r[7];
            if constexpr (PAYOFF == kPayAsian) {
                suma = __uint_as_float(r[8]);
                sumb = __uint_as_float(r[9]);
                fa = (float)ya;
                fb = (float)yb;
            }
            if constexpr (PAYOFF == kPayUpAndOut) {
                mxa = get_f64(r + 8);
                mxb = get_f64(r + 10);
            }
        };
        auto block_end = [&](uint32_t s, PairStream& st) {
            fold(s);
            if constexpr (GROUP) {
                const float vk = ANTI ? fminf(va, vb) : va;
                move(group_dest(g, vk < d.qe_vlow, (blockIdx.x + (s >> 3)) & 3u), st);
            }
        };
        philox_steps<true, JUMPS, kDefer, kMixed>(d, src.st, step, &jsum, block_end);
        if constexpr (GROUP) {  // every pair back to its home lane, its last partial block folded
            fold(N);
            PairStream st = src.st;
            __syncthreads();
            move(home, st);
        }
    } else {
        for (uint32_t i = 0; i < N; ++i) {
            double za, zb, j;
            src.get(i, za, zb, j);
            step((float)za, (float)zb, (float)(j * kLog2e), 0u, i);
            if ((i & 7u) == 7u) fold(i);
        }
    }
    fold(N);  // the last partial block (empty when 8 divides N: pa = 0, ma = -inf)
    if (PAYOFF == kPayEuropean) {
        // the Euler constant N a0 and the jumps (Euler: summed in the steps;
        // QE: only those drawn after the loop)
        double c = (double)jsum;
        if (SCHEME == kSchemeEuler) c += (double)N * (double)d.eu_a0;
        ya += c;
        yb += c;
    }

    const double nan = __builtin_nan("");
    auto finish = [&](PathState& e, double y, float v, float sum, double mx) {
        e.ln_s = d.ln_spot + y * kLn2;
        e.v = (double)v;
        e.sum = PAYOFF == kPayAsian ? d.spot * (double)sum : nan;
        e.max = nan;
        e.ko = 0;
        e.near = 0;
        if (PAYOFF == kPayAsian) {
            e.x = d.spot * (double)sum / (double)N;
        } else {
            e.x = d.spot * exp2(y);
            if (PAYOFF == kPayUpAndOut) {
                // barrier test in FP64 on the FP64 log-price maximum
                e.ko = (d.spot >= d.barrier) || (mx >= d.lb2_barrier);
                e.max = fmax(d.spot, d.spot * exp2(mx));
                e.near = fabs(mx - d.lb2_barrier) < kBarrierMargin;
            }
        }
    };
    finish(ea, ya, va, suma, mxa);
    if (anti) finish(eb, yb, vb, sumb, mxb);
}

// GROUP: as sim_pair_f32_loop (lds: the block's LDS).
template <int SCHEME, bool JUMPS, int PAYOFF, bool GROUP = false, class Src>
__device__ __forceinline__ void sim_pair_f32(const McDesc& d, const Src& src, bool anti, PathState& ea,
                                             PathState& eb, void* lds = nullptr) {
    if (SCHEME == kSchemeQE && !anti)
        sim_pair_f32_loop<SCHEME, JUMPS, PAYOFF, false, GROUP>(d, src, false, ea, eb, lds);
    else
        sim_pair_f32_loop<SCHEME, JUMPS, PAYOFF, true, GROUP>(d, src, anti, ea, eb, lds);
    if constexpr (PAYOFF == kPayUpAndOut) {
        // With the call's FP64 barrier flag, the barrier decision near B is
        // FP64's: the pair is simulated again on the same draws by the FP64
        // path of the same mode (Philox or replay), and a sibling whose
        // maximum came within kBarrierMargin of the barrier takes that path's
        // state, knock-out and terminal value.  Rare per path (about 2
        // margins times the density of the maximum at B), but the wave that
        // runs it becomes the kernel's critical path, hence opt-in.  The
        // kernels' occupancy is pinned so that this cold code does not raise
        // the loop's register budget.
        const bool na = ea.near, nb = anti && eb.near;
        if ((d.flags & 1u) && (na || nb)) {
            PathState fa, fb;
            sim_pair_f64<SCHEME, JUMPS, PAYOFF>(d, src, anti, fa, fb);
            if (na) ea = fa;
            if (nb) eb = fb;
        }
    }
}

// ------------------------------------------------------------ reductions
__device__ __forceinline__ void wave_sum2(double& s, double& q) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        s += __shfl_xor(s, m, 64);
        q += __shfl_xor(q, m, 64);
    }
}

// Sum of the n_tiles partial-sum slots slots[base + t] of one (request,
// strike), t in tile order: thread i takes tiles i, i + kBlock, ...
// sequentially (coalesced across the block), then the xor butterfly over the
// wave, then waves 0..3 in order.  Every reduction of a request's slots goes
// through this one function, so the order (and the bits) never depend on
// which kernel reduces.  The result is valid in thread 0.
__device__ __forceinline__ double2 block_sum_slots(const double2* slots, uint64_t base, uint32_t n_tiles,
                                                   double2* w) {
    double s = 0.0, q = 0.0;
    for (uint32_t t = threadIdx.x; t < n_tiles; t += kBlock) {
        const double2 v = slots[base + t];
        s += v.x;
        q += v.y;
    }
    wave_sum2(s, q);
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = make_double2(s, q);
    __syncthreads();
    return make_double2(((w[0].x + w[1].x) + w[2].x) + w[3].x, ((w[0].y + w[1].y) + w[2].y) + w[3].y);
}

// Per-path outputs (replay mode only).
struct PathSink {
    double* payoff;  // [pair][2]
    double* state;   // [pair][2][4]
};

// LDS of a tile block (double2 words): the FP64 strike loop keeps one
// (sum, sum_sq) per (strike, wave); the FP32 strike loop transposes the
// lane sums of kEpiStrikes strikes through rows of kEpiRow float2 (one per
// thread, padded by one against bank conflicts), then keeps one FP64 pair
// per (strike of the chunk, wave) at kEpiWsumOff.
constexpr int kEpiStrikes = 8;
constexpr int kEpiRow = kBlock + 1;
constexpr int kEpiWsumOff = (kEpiStrikes * kEpiRow * 8 + 15) / 16;
constexpr int kTileLds = kMaxStrikes * 4 > kEpiWsumOff + kEpiStrikes * 4 ? kMaxStrikes * 4
                                                                          : kEpiWsumOff + kEpiStrikes * 4;
static_assert(kBlock == 256 && kEpiStrikes == 8, "thread t sums threads 8 (t >> 3) .. +7 of row t & 7");
static_assert(sizeof(GroupLds) <= sizeof(double2) * kTileLds, "the regrouping's rows fit the tile LDS");

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
    // A tile never straddles a multiple of 2^32 pairs (kBlock * PPT divides
    // 2^32), so the high word of every pair index in it is the tile's: a
    // uniform value, which lets the first Philox round run on the scalar unit.
    const uint32_t pair_hi = (uint32_t)(base >> 32);
    // ea/eb/valid are only ever indexed with compile-time constants so they
    // stay in registers; the simulation loop over j is kept rolled (one copy
    // of the step code) and hands its result over through selects.
    PathState ea[PPT], eb[PPT];
    bool valid[PPT];
    // QE in Philox mode regroups the block's pairs by variance (see
    // group_dest), so every thread steps, also those past the last pair of a
    // ragged tile: their draws are Philox words of pair indices no request
    // owns, and their results are dropped.
    using Src = decltype(make_src(d, uint64_t(0), uint32_t(0)));
    constexpr bool kGroup = SCHEME == kSchemeQE && Src::kPhilox && !(JUMPS && PAYOFF == kPayEuropean);
#pragma nounroll
    for (int j = 0; j < PPT; ++j) {
        const uint64_t pair = base + (uint64_t)j * kBlock + (uint64_t)tid;
        const bool ok = pair < d.n_pairs;
        PathState a, b;
        if (ok || kGroup) {
            const auto src = make_src(d, pair, pair_hi);
            if constexpr (sizeof(Real) == 8)
                sim_pair_f64<SCHEME, JUMPS, PAYOFF, kGroup>(d, src, anti, a, b, lds);
            else
                sim_pair_f32<SCHEME, JUMPS, PAYOFF, kGroup>(d, src, anti, a, b, lds);
        }
        if (ok) {
            if (sink.payoff) {
                const double k0 = strikes[d.strike_off];
                sink.payoff[pair * 2] = payoff_value(d, a, k0);
                sink.payoff[pair * 2 + 1] = anti ? payoff_value(d, b, k0) : __builtin_nan("");
            }
            if (sink.state) {
                double* o = sink.state + pair * 8;
                o[0] = a.ln_s; o[1] = a.v; o[2] = a.sum; o[3] = a.max;
                if (anti) { o[4] = b.ln_s; o[5] = b.v; o[6] = b.sum; o[7] = b.max; }
            }
        }
#pragma unroll
        for (int jj = 0; jj < PPT; ++jj) {
            if (jj == j) {
                ea[jj] = a;
                eb[jj] = b;
                valid[jj] = ok;
            }
        }
    }
    // the strike loop reuses the LDS the regrouping used: every thread has
    // taken its pair back first
    if constexpr (kGroup) __syncthreads();

    const int lane = tid & 63, wave = tid >> 6;
    const uint32_t ns = d.n_strikes;
    // the lane's payoffs are summed undiscounted and the sums discounted once
    // (disc and disc^2): sum and sum of squares of disc * payoff up to rounding
    const double sgn = d.opt_put ? -1.0 : 1.0, disc2 = d.disc * d.disc;
    if constexpr (sizeof(Real) == 4) {
        // FP32 path: each lane's own payoffs (at most 2 PPT) and their sums
        // in FP32; every sum over lanes, waves and tiles in FP64.  A path that
        // does not exist or is knocked out enters as sgn x = -inf, whose
        // intrinsic value is 0; the knocked-out ones add the rebate instead,
        // the same for every strike.
        constexpr int M = 2 * PPT;
        float sx[M];
        float nko = 0.0f;
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            const bool oka = valid[j], okb = valid[j] && anti;
            sx[2 * j] = oka && !ea[j].ko ? (float)(sgn * ea[j].x) : -__builtin_huge_valf();
            sx[2 * j + 1] = okb && !eb[j].ko ? (float)(sgn * eb[j].x) : -__builtin_huge_valf();
            nko += (oka && ea[j].ko ? 1.0f : 0.0f) + (okb && eb[j].ko ? 1.0f : 0.0f);
        }
        const float cs = nko * (float)d.rebate, cq = cs * (float)d.rebate;
        // Strikes in chunks of kEpiStrikes: the block's lane sums go through
        // LDS transposed (row c of the chunk holds every thread's sums for
        // strike c), thread t then adds the 8 threads 8 (t >> 3) .. +7 of row
        // t & 7 in FP64, the xor butterfly over the wave's 8 groups and the
        // 4 waves in order complete the tile's sum, and the sums are
        // discounted once.  The order depends on the request alone.
        float2* rows = reinterpret_cast<float2*>(lds);
        double2* wsum = lds + kEpiWsumOff;
        const int c = tid & (kEpiStrikes - 1), g = tid >> 3;
        for (uint32_t k0 = 0; k0 < ns; k0 += kEpiStrikes) {
#pragma unroll
            for (int cc = 0; cc < kEpiStrikes; ++cc) {
                float s = 0.0f, q = 0.0f;
                if (k0 + cc < ns) {
                    const float nk = (float)(-sgn * strikes[d.strike_off + k0 + cc]);
                    s = cs;
                    q = cq;
#pragma unroll
                    for (int i = 0; i < M; ++i) {
                        const float p = fmaxf(sx[i] + nk, 0.0f);
                        s += p;
                        q = fmaf(p, p, q);
                    }
                }
                rows[cc * kEpiRow + tid] = make_float2(s, q);
            }
            __syncthreads();
            double s = 0.0, q = 0.0;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const float2 v = rows[c * kEpiRow + g * 8 + i];
                s += (double)v.x;
                q += (double)v.y;
            }
#pragma unroll
            for (int m = 8; m < 64; m <<= 1) {
                s += __shfl_xor(s, m, 64);
                q += __shfl_xor(q, m, 64);
            }
            if (lane < kEpiStrikes) wsum[lane * 4 + wave] = make_double2(s, q);
            __syncthreads();
            if (tid < kEpiStrikes && k0 + tid < ns) {
                const double2 a = wsum[tid * 4 + 0], b = wsum[tid * 4 + 1], e = wsum[tid * 4 + 2],
                              f = wsum[tid * 4 + 3];
                slots[d.slot_off + (uint64_t)(k0 + tid) * d.n_tiles + tile] =
                    make_double2((((a.x + b.x) + e.x) + f.x) * d.disc, (((a.y + b.y) + e.y) + f.y) * disc2);
            }
        }
    } else {
        for (uint32_t k = 0; k < ns; ++k) {
            const double nk = -sgn * strikes[d.strike_off + k];
            double s = 0.0, q = 0.0;
#pragma unroll
            for (int j = 0; j < PPT; ++j) {
                if (valid[j]) {
                    const double pa = payoff_undiscounted(d, ea[j], sgn, nk);
                    s += pa;
                    q += pa * pa;
                    if (anti) {
                        const double pb = payoff_undiscounted(d, eb[j], sgn, nk);
                        s += pb;
                        q += pb * pb;
                    }
                }
            }
            s *= d.disc;
            q *= disc2;
            wave_sum2(s, q);
            if (lane == 0) lds[k * 4 + wave] = make_double2(s, q);
        }
        __syncthreads();
        for (uint32_t k = tid; k < ns; k += kBlock) {
            const double2 a = lds[k * 4 + 0], b = lds[k * 4 + 1], c = lds[k * 4 + 2], e = lds[k * 4 + 3];
            slots[d.slot_off + (uint64_t)k * d.n_tiles + tile] =
                make_double2(((a.x + b.x) + c.x) + e.x, ((a.y + b.y) + c.y) + e.y);
        }
    }
}

}  // namespace
}  // namespace ope
