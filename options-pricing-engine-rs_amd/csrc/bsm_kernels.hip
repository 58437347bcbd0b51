// Black-Scholes-Merton chains (bsm.rs:11-54) and implied volatility
// (iv.rs:14-117) on the GPU: one lane per option, FP64, the reference's
// expressions in its evaluation order (built with -ffp-contract=off).
//
// The price-and-Greeks kernel reads a 56-byte contract and writes a 64-byte
// result per option with a few dozen FP64 operations in between: it is bound
// by HBM bandwidth, and stages its results through LDS.  The
// implied-volatility kernel re-prices each option 3-15 times (guess check,
// Halley steps, or the bisection fallback) and is bound by FP64 issue; its
// lanes diverge only in their iteration counts.
#include <hip/hip_runtime.h>

#include "../../include/ope_cf.h"

namespace ope {
namespace {

constexpr double kSqrt2Inv = 0.7071067811865476;    // std::f64::consts::FRAC_1_SQRT_2
constexpr double kSqrt2PiInv = 0.3989422804014327;  // math.rs:6

__device__ __forceinline__ double ncdf(double x) { return 0.5 * erfc(-x * kSqrt2Inv); }  // math.rs:16-19
__device__ __forceinline__ double npdf(double x) { return kSqrt2PiInv * exp(-0.5 * x * x); }  // math.rs:8-13

// bsm.rs:49-54
__device__ double bsm_price(double s, double k, double t, double r, double q, double v, double phi) {
    const double vt = v * sqrt(t);
    const double d1 = (log(s / k) + (r - q + 0.5 * v * v) * t) / vt;
    const double d2 = d1 - vt;
    return phi * (s * exp(-q * t) * ncdf(phi * d1) - k * exp(-r * t) * ncdf(phi * d2));
}

// bsm.rs:11-45 (theta keeps the reference's sign of the carry term)
__device__ ope_pricing_result bsm_greeks(double s, double k, double t, double r, double q, double v, double phi) {
    const double vt = v * sqrt(t);
    const double d1 = (log(s / k) + (r - q + 0.5 * v * v) * t) / vt;
    const double d2 = d1 - vt;
    const double nd1 = ncdf(phi * d1), nd2 = ncdf(phi * d2), npd1 = npdf(d1);
    const double eq = exp(-q * t), er = exp(-r * t);
    const double seq = s * eq, ker = k * er, sqt = sqrt(t);
    ope_pricing_result g;
    g.price = phi * (seq * nd1 - ker * nd2);
    g.delta = phi * eq * nd1;
    g.gamma = eq * npd1 / (s * v * sqt);
    g.vega = seq * npd1 * sqt;
    g.theta = -seq * npd1 * (v * sqt) / (2.0 * t) - phi * (q * seq * nd1 - r * ker * nd2);
    g.rho = phi * k * t * er * nd2;
    g.vanna = -eq * npd1 * d2 / v;
    g.volga = g.vega * d1 * d2 / v;
    return g;
}

constexpr int kBsmBlock = 256;
constexpr int kResultWords = (int)(sizeof(ope_pricing_result) / 8);     // 8
static_assert(sizeof(ope_option_contract) == 56 && sizeof(ope_pricing_result) == 64, "record layouts");

// Each thread reads its contract directly (the 56-byte records of a wave
// are contiguous, so the L2 sees whole lines), but the 64-byte results go
// through LDS and leave as one contiguous run of 8-byte words per block
// (consecutive lanes, consecutive addresses): per-lane 64-byte-strided stores
// ran at 4.1 TB/s, the staged ones at 5.3 TB/s, and staging the loads as
// well was slower (3.7 TB/s).  The staging row is padded to 9 words to
// spread the per-thread writes over the LDS banks.
__global__ __launch_bounds__(kBsmBlock) void k_bsm(const ope_option_contract* __restrict__ c, uint32_t n,
                                                   double* __restrict__ price, ope_pricing_result* __restrict__ greeks) {
    __shared__ double s_out[kBsmBlock * (kResultWords + 1)];
    const uint32_t base = blockIdx.x * (uint32_t)kBsmBlock;
    const uint32_t cnt = min((uint32_t)kBsmBlock, n - base);
    const uint32_t tid = threadIdx.x;
    if (tid < cnt) {
        const ope_option_contract o = c[base + tid];
        const double phi = o.opt_type == OPE_CALL ? 1.0 : -1.0;
        if (greeks) {
            const ope_pricing_result g = bsm_greeks(o.spot, o.strike, o.expiry, o.rate, o.div_yield, o.vol, phi);
            const double* gv = reinterpret_cast<const double*>(&g);
#pragma unroll
            for (int f = 0; f < kResultWords; ++f) s_out[tid * (kResultWords + 1) + f] = gv[f];
        } else {
            price[base + tid] = bsm_price(o.spot, o.strike, o.expiry, o.rate, o.div_yield, o.vol, phi);
        }
    }
    if (greeks) {
        __syncthreads();
        double* dst = reinterpret_cast<double*>(greeks + base);
        for (uint32_t j = tid; j < cnt * kResultWords; j += kBsmBlock)
            dst[j] = s_out[(j / kResultWords) * (kResultWords + 1) + (j % kResultWords)];
    }
}

// iv.rs:14-117: bounds, Brenner-Subrahmanyam guess checked against 20 % of
// the price (else bisection on [1e-4, 10]), then at most 10 Halley steps.
__device__ bool implied_vol(const ope_option_contract& o, double price, double& vol) {
    const double s = o.spot, k = o.strike, t = o.expiry, r = o.rate, q = o.div_yield;
    const double phi = o.opt_type == OPE_CALL ? 1.0 : -1.0;
    if (price <= 0.0) return false;  // price_in_bounds (iv.rs:22-32)
    const double er = exp(-r * t), eq = exp(-q * t);
    const double intrinsic = phi > 0.0 ? fmax(s * eq - k * er, 0.0) : fmax(k * er - s * eq, 0.0);
    if (!(price >= intrinsic)) return false;
    const double fwd = s * eq / er;  // initial_guess (iv.rs:37-50)
    const double bs = price * sqrt(2.0 * 3.141592653589793 / t) / (fwd * er);
    const double x = log(fwd / k);
    double v = bs / fmax(exp(-0.5 * x * x), 0.01);
    v = v < 0.001 ? 0.001 : (v > 10.0 ? 10.0 : v);
    if (!(fabs(bsm_price(s, k, t, r, q, v, phi) - price) < 0.2 * price)) {
        double lo = 1e-4, hi = 10.0;  // bisect (iv.rs:53-65)
        if ((bsm_price(s, k, t, r, q, lo, phi) - price) * (bsm_price(s, k, t, r, q, hi, phi) - price) > 0.0)
            return false;
        bool done = false;
        for (int i = 0; i < 60 && !done; ++i) {
            const double mid = 0.5 * (lo + hi);
            if (bsm_price(s, k, t, r, q, mid, phi) - price < 0.0)
                lo = mid;
            else
                hi = mid;
            if (hi - lo < 1e-9) {
                v = mid;
                done = true;
            }
        }
        if (!done) v = 0.5 * (lo + hi);
    }
    for (int it = 0; it < 10; ++it) {  // halley_solve (iv.rs:68-90), vega_and_volga (93-103)
        const double err = bsm_price(s, k, t, r, q, v, phi) - price;
        if (fabs(err) < 1e-10) {
            vol = v;
            return true;
        }
        const double vt = v * sqrt(t);
        const double d1 = (log(s / k) + (r - q + 0.5 * v * v) * t) / vt;
        const double d2 = d1 - vt;
        const double vega = s * eq * npdf(d1) * sqrt(t);
        const double volga = vega * d1 * d2 / v;
        if (fabs(vega) < 1e-14) return false;
        double denom = 1.0 - err * volga / (2.0 * vega * vega);
        denom = denom < 0.5 ? 0.5 : (denom > 2.0 ? 2.0 : denom);
        v -= err / (vega * denom);
        if (v <= 0.0) v = 1e-8;
    }
    if (fabs(bsm_price(s, k, t, r, q, v, phi) - price) < 1e-6) {
        vol = v;
        return true;
    }
    return false;
}

__global__ __launch_bounds__(256) void k_iv(const ope_option_contract* __restrict__ c,
                                            const double* __restrict__ mkt, uint32_t n, double* __restrict__ vol,
                                            uint8_t* __restrict__ found) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    double v = __builtin_nan("");
    const bool ok = implied_vol(c[i], mkt[i], v);
    vol[i] = ok ? v : __builtin_nan("");
    found[i] = ok ? 1 : 0;
}

}  // namespace

hipError_t launch_bsm(const ope_option_contract* c, uint32_t n, double* price, ope_pricing_result* greeks,
                      hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_bsm, dim3((n + kBsmBlock - 1) / kBsmBlock), dim3(kBsmBlock), 0, s, c, n, price, greeks);
    return hipGetLastError();
}

hipError_t launch_iv(const ope_option_contract* c, const double* mkt, uint32_t n, double* vol, uint8_t* found,
                     hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_iv, dim3((n + 255) / 256), dim3(256), 0, s, c, mkt, n, vol, found);
    return hipGetLastError();
}

}  // namespace ope
