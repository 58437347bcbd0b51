// Internal interface of the host side of the engine, shared by the C ABI
// (capi.cpp) and the serving loop (server.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/ope_mc.h"
#include "mc_desc.h"

namespace ope {
namespace host {

extern thread_local std::string g_last_error;

// Records a formatted message for ope_last_error_message() and returns code.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

// Everything one pricing simulation needs.
struct Request {
    double spot, expiry, rate, div_yield;
    ope_heston_params hp;
    bool has_jump_params;  // mc_bates (drift carries the jump compensator)
    double lambda, mu_j, sigma_j;
    ope_payoff payoff;
    ope_mc_config cfg;
};

// Validates a request and fills its descriptor (everything except the
// planning fields: ppt, n_tiles, n_strikes, strike_off, out_off, slot_off).
int make_desc(const Request& r, McDesc& d);

Request make_request(double spot, double expiry, double rate, double div_yield, const ope_heston_params* hp,
                     bool bates, double lambda, double mu_j, double sigma_j, const ope_payoff* payoff,
                     const ope_mc_config* cfg);

// OPE_OK when a HIP device is usable (initialises the shared context once).
int device_available();

// The shared context's stream and timing events, as a call sees them.
struct ContextView {
    hipStream_t stream;
    hipEvent_t ev0, evk, ev1;  // before the pricing launch, after it, after the reduction
};

// Locks the shared context (every entry point is serialised by it),
// initialises it on first use, and runs body(arg, view) under the lock.
int with_context(int (*body)(void* arg, const ContextView& view), void* arg);

// Checks an options struct (precision and RNG mode).
int check_options(const ope_mc_options* o);

// A single request's tiling as every entry point plans it: pairs per thread
// (4 from 2^22 pairs, else 1), tiles of kBlock * ppt pairs, one strike.
// Returns OPE_OK or OPE_ERR_INVALID_ARGUMENT for too many tiles.
int set_single_tiling(McDesc& d);

// Price and standard error from (sum, sum_sq), exactly as mc.rs:198-204.
void finish(const McDesc& d, double2 sums, ope_mc_result* out, ope_mc_stats* st);

}  // namespace host
}  // namespace ope

#define HIP_TRY(expr)                                                                                   \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            return fail(e_ == hipErrorOutOfMemory ? OPE_ERR_OUT_OF_MEMORY : OPE_ERR_HIP, "%s: %s (%s:%d)", \
                        #expr, hipGetErrorString(e_), __FILE__, __LINE__);                             \
    } while (0)

