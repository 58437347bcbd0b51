// ope_mc_bates_noalloc: ope_mc_bates_ex without any allocation, for callers
// that cannot call hipMalloc (code under HIP graph capture, tight serving
// loops, harnesses that forbid it).
//
// The call runs as one direct launch (mc_direct.hip): the descriptor and the
// strike travel in the kernel arguments, and the device buffers it needs live
// in `scratch`, a file-scope device array of kScratchBytes: the launch's
// completion counter, the reduced sums and one (sum, sum_sq) slot per tile.
// At the first call on a device (with that device current) hipGetSymbolAddress
// gives the array's device address, kept per device, and a one-thread kernel
// clears the counter; host code passes only that address to kernels and
// copies, never the array's name.  After that, each launch's last block clears
// the counter again; every slot is written by its tile before the reduction
// reads it.
//
// ope_mc_bates_ex runs a request of up to 16 strikes as the same direct
// launch (same tiling, same kernel, same tile-order reduction), so the result
// is bit for bit ope_mc_bates_ex's.  Calls are serialised by the library's
// context lock (calls that overlapped on one GPU would race on the array); the
// array keeps the last call's contents.
//
// Limit: at most kMaxTiles tiles, i.e. 2^24 antithetic pairs (2^25 paths).
#include <hip/hip_runtime.h>

#include <cstring>

#include "engine.h"
#include "mc_launch.h"

namespace ope {
namespace {

constexpr uint32_t kMaxTiles = 16384;
constexpr size_t kHeaderBytes = 256;  // completion counter, reduced sums
constexpr size_t kOffDone = 0, kOffOut = 16;
constexpr size_t kOffSlots = kHeaderBytes;
constexpr size_t kScratchBytes = kOffSlots + (size_t)kMaxTiles * sizeof(double2);
constexpr int kMaxDevices = 64;

static_assert(kOffDone + sizeof(uint32_t) <= kOffOut && kOffOut + sizeof(double2) <= kHeaderBytes, "header layout");
static_assert(kOffSlots % 256 == 0, "slot alignment");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on each device (null until that device's first call).
char* g_scratch[kMaxDevices] = {};

__global__ void k_clear_counter(uint32_t* done) { *done = 0u; }

struct Call {
    McDesc d;
    double strike;
    int fp32;
    double2 sums;
    float ms;
};

// Runs under the context lock (host::with_context).
int run(void* arg, const host::ContextView& v) {
    using host::fail;
    Call& c = *static_cast<Call*>(arg);
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(OPE_ERR_NO_DEVICE, "device %d beyond the supported range", dev);
    hipStream_t s = v.stream;
    if (!g_scratch[dev]) {  // set-up, once per device, with that device current
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        hipLaunchKernelGGL(k_clear_counter, dim3(1), dim3(1), 0, s,
                           reinterpret_cast<uint32_t*>(static_cast<char*>(p) + kOffDone));
        HIP_TRY(hipGetLastError());
        g_scratch[dev] = static_cast<char*>(p);
    }
    char* base = g_scratch[dev];
    double2* d_out = reinterpret_cast<double2*>(base + kOffOut);

    DirectLaunch L;
    std::memset(&L, 0, sizeof L);
    L.args.d = c.d;
    L.args.strikes[0] = c.strike;
    L.slots = reinterpret_cast<double2*>(base + kOffSlots);
    L.done = reinterpret_cast<uint32_t*>(base + kOffDone);
    L.out = d_out;
    L.fp32 = c.fp32;
    HIP_TRY(hipEventRecord(v.ev0, s));
    HIP_TRY(launch_direct(L, s));
    HIP_TRY(hipEventRecord(v.ev1, s));
    HIP_TRY(hipMemcpyAsync(&c.sums, d_out, sizeof(double2), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    HIP_TRY(hipEventElapsedTime(&c.ms, v.ev0, v.ev1));
    return OPE_OK;
}

}  // namespace
}  // namespace ope

using namespace ope;
using namespace ope::host;

extern "C" int ope_mc_bates_noalloc(double spot, double expiry, double rate, double div_yield,
                                    const ope_heston_params* params, double lambda, double mu_j, double sigma_j,
                                    const ope_payoff* payoff, const ope_mc_config* cfg, const ope_mc_options* opts,
                                    ope_mc_result* out) {
    if (!params || !payoff || !cfg || !out) return fail(OPE_ERR_INVALID_ARGUMENT, "null argument");
    int rc = check_options(opts);
    if (rc) return rc;
    if (opts && (opts->rng == OPE_RNG_REPLAY || opts->path_payoffs || opts->path_states))
        return fail(OPE_ERR_INVALID_ARGUMENT, "the no-alloc entry point takes Philox draws only, without per-path outputs");
    Call c;
    std::memset(&c, 0, sizeof c);
    const Request r = make_request(spot, expiry, rate, div_yield, params, true, lambda, mu_j, sigma_j, payoff, cfg);
    if ((rc = make_desc(r, c.d))) return rc;
    c.d.flags = opts ? opts->flags : 0u;
    if ((rc = set_single_tiling(c.d))) return rc;
    if (c.d.n_tiles > kMaxTiles)
        return fail(OPE_ERR_INVALID_ARGUMENT,
                    "%llu pairs need %u tiles; the no-alloc entry point's scratch holds %u (2^24 pairs)",
                    (unsigned long long)c.d.n_pairs, c.d.n_tiles, kMaxTiles);
    c.strike = payoff->strike;
    c.fp32 = opts && opts->precision == OPE_PRECISION_FP32_PATH;
    if ((rc = with_context(run, &c))) return rc;
    finish(c.d, c.sums, out, opts ? opts->stats : nullptr);
    if (opts && opts->device_ms) *opts->device_ms = c.ms;
    if (opts && opts->kernel_ms) *opts->kernel_ms = c.ms;
    return OPE_OK;
}
