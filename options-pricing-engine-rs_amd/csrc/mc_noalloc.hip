// ope_mc_bates_noalloc: ope_mc_bates_ex without any allocation, for callers
// that cannot call hipMalloc (code under HIP graph capture, tight serving
// loops, harnesses that forbid it).
//
// Every device buffer of the call lives in `scratch`, a file-scope device
// array of kScratchBytes: the request's descriptor, strike and reduction job,
// the tile list, one (sum, sum_sq) slot per tile, and the reduced sums.  At
// the first call on a device (with that device current) hipGetSymbolAddress
// gives the array's device address, kept per device; host code passes only
// that address to kernels and copies, never the array's name.
//
// The call plans, launches and reduces exactly as ope_mc_bates_ex does for
// one request (same tiling, same specialised kernel, same tile-order
// reduction), so its result is bit-for-bit that of ope_mc_bates_ex.  Nothing
// in the array has to start at zero: every slot is written by its tile
// before the reduction reads it.  Calls are serialised by the library's
// context lock; the array keeps the last call's contents.
//
// Limit: at most kMaxTiles tiles, i.e. 2^24 antithetic pairs (2^25 paths).
#include <hip/hip_runtime.h>

#include <cstring>

#include "engine.h"
#include "mc_launch.h"

namespace ope {
namespace {

constexpr uint32_t kMaxTiles = 16384;
constexpr size_t kHeaderBytes = 1024;  // descriptor, strike, job, reduced sums
constexpr size_t kOffDesc = 0, kOffStrike = 512, kOffJob = 528, kOffOut = 544;
constexpr size_t kOffTiles = kHeaderBytes;
constexpr size_t kOffSlots = kOffTiles + (size_t)kMaxTiles * sizeof(uint2);
constexpr size_t kScratchBytes = kOffSlots + (size_t)kMaxTiles * sizeof(double2);
constexpr int kMaxDevices = 64;

static_assert(sizeof(McDesc) <= kOffStrike, "descriptor fits its header field");
static_assert(kOffStrike + sizeof(double) <= kOffJob && kOffJob + sizeof(uint2) <= kOffOut, "header layout");
static_assert(kOffOut + sizeof(double2) <= kHeaderBytes && kOffSlots % 256 == 0, "header layout");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on each device (null until that device's first call).
char* g_scratch[kMaxDevices] = {};

// The single request's tile list: tile t of request 0.
__global__ void k_single_tiles(uint2* tiles, uint32_t n) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) tiles[t] = make_uint2(0u, t);
}

struct Call {
    McDesc d;
    double strike;
    int fp32;
    double2 sums;
    float ms, kms;
};

// Runs under the context lock (host::with_context).
int run(void* arg, const host::ContextView& v) {
    using host::fail;
    Call& c = *static_cast<Call*>(arg);
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(OPE_ERR_NO_DEVICE, "device %d beyond the supported range", dev);
    if (!g_scratch[dev]) {  // set-up, once per device, with that device current
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[dev] = static_cast<char*>(p);
    }
    char* base = g_scratch[dev];
    McDesc* d_desc = reinterpret_cast<McDesc*>(base + kOffDesc);
    double* d_strike = reinterpret_cast<double*>(base + kOffStrike);
    uint2* d_job = reinterpret_cast<uint2*>(base + kOffJob);
    double2* d_out = reinterpret_cast<double2*>(base + kOffOut);
    uint2* d_tiles = reinterpret_cast<uint2*>(base + kOffTiles);
    double2* d_slots = reinterpret_cast<double2*>(base + kOffSlots);

    // descriptor, strike and job in one copy from the caller's stack
    struct alignas(16) Header {
        unsigned char bytes[kOffOut];
    } h;
    std::memset(&h, 0, sizeof h);
    std::memcpy(h.bytes + kOffDesc, &c.d, sizeof(McDesc));
    std::memcpy(h.bytes + kOffStrike, &c.strike, sizeof(double));
    const uint2 job = make_uint2(0u, 0u);
    std::memcpy(h.bytes + kOffJob, &job, sizeof job);
    hipStream_t s = v.stream;
    HIP_TRY(hipMemcpyAsync(base, h.bytes, sizeof h.bytes, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_single_tiles, dim3((c.d.n_tiles + 255) / 256), dim3(256), 0, s, d_tiles, c.d.n_tiles);
    HIP_TRY(hipGetLastError());

    TileLaunch L;
    L.descs = d_desc;
    L.tiles = d_tiles;
    L.n_tiles = c.d.n_tiles;
    L.strikes = d_strike;
    L.slots = d_slots;
    L.fp32 = c.fp32;
    L.fixed = 1;
    L.code = config_code(c.d);
    HIP_TRY(hipEventRecord(v.ev0, s));
    HIP_TRY(launch_tiles(L, s));
    HIP_TRY(hipEventRecord(v.evk, s));
    HIP_TRY(launch_reduce(d_desc, d_job, 1, d_slots, d_out, s));
    HIP_TRY(hipEventRecord(v.ev1, s));
    HIP_TRY(hipMemcpyAsync(&c.sums, d_out, sizeof(double2), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    HIP_TRY(hipEventElapsedTime(&c.ms, v.ev0, v.ev1));
    HIP_TRY(hipEventElapsedTime(&c.kms, v.ev0, v.evk));
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
    if (opts && opts->kernel_ms) *opts->kernel_ms = c.kms;
    return OPE_OK;
}
