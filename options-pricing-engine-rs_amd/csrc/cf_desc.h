// Descriptor of one analytic (characteristic-function) price, shared by the
// host (cf.cpp) and the kernel (cf_kernels.hip).
#pragma once

#include <stdint.h>

namespace ope {

struct CfDesc {
    double spot, strike, expiry, rate, div_yield;
    double v0, kappa, theta, sigma, rho;
    double lambda, mu_j, sigma_j;
    int32_t bates;  // 1: Bates (Heston CF x jump CF), 0: Heston
    int32_t put;    // 1: put by parity from the call
    int32_t deriv;  // Bates in a derivative launch: also d(price)/d(lambda, mu_j, sigma_j)
    int32_t pad_;
};

// Adaptive Gauss-Kronrod limits of the reference (heston.rs:193-195, 228).
constexpr int kCfMaxPanels = 200;
constexpr double kCfTol = 1e-8;

// Diagnostics per option: panels used by the P1 and P2 integrals.
struct CfStats {
    uint32_t panels_p1;
    uint32_t panels_p2;
};

}  // namespace ope
