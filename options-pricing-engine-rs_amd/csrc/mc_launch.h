// Host-side launch interface between the C ABI (capi.cpp) and the kernel
// translation units (mc_kernels.hip, mc_replay.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mc_desc.h"

namespace ope {

// Configuration code of a request: which specialised path-simulation body
// runs it.  0..23 = ((scheme * 2 + jumps) * 3 + payoff) * 2 + (ppt == 4).
constexpr uint32_t kNumCodes = 24;
__host__ __device__ inline uint32_t config_code(const McDesc& d) {
    return ((d.scheme * 2u + d.jumps) * 3u + d.payoff) * 2u + (d.ppt == 4 ? 1u : 0u);
}

struct TileLaunch {
    const McDesc* descs;   // device
    const uint2* tiles;    // device: (request, tile within request)
    uint32_t n_tiles;
    const double* strikes; // device
    double2* slots;        // device: (sum, sum_sq) per (tile, strike)
    int fp32;              // 0: FP64 path, 1: FP32 path
    int fixed;             // 1: every request has config code `code`
    uint32_t code;
};

// A direct launch: one request whose descriptor and strikes travel in the
// kernel arguments (no upload), tile = block index, and the block that
// finishes the last tile sums every strike's slots in tile order into `out`
// (block_sum_slots, as k_reduce: the same bits), then clears `done`.
constexpr uint32_t kDirectStrikes = 16;
struct DirectArgs {
    McDesc d;                        // strike_off = 0: the strikes below
    double strikes[kDirectStrikes];
};
struct DirectLaunch {
    DirectArgs args;
    double2* slots;        // device
    uint32_t* done;        // device, 0 before the launch; the last block resets it
    double2* out;          // (sum, sum_sq) per strike at args.d.out_off; host-coherent memory
    int fp32;
};

struct ReplayLaunch {
    const McDesc* desc;    // device, one request, ppt == 1
    uint32_t n_tiles;
    const double* za;      // device [n_steps][n_pairs]
    const double* zb;
    const double* jump;    // may be null
    const double* strikes;
    double2* slots;
    double* path_payoff;   // device [n_pairs][2] or null
    double* path_state;    // device [n_pairs][2][4] or null
    int fp32;
    uint32_t code;
};

hipError_t launch_tiles(const TileLaunch& L, hipStream_t s);
hipError_t launch_reduce(const McDesc* descs, const uint2* jobs, uint32_t n_jobs, const double2* slots,
                         double2* out, hipStream_t s);
hipError_t launch_replay(const ReplayLaunch& L, hipStream_t s);
hipError_t launch_direct(const DirectLaunch& L, hipStream_t s);
hipError_t launch_debug_philox(uint64_t seed, const uint4* ctr, uint4* out, uint32_t n, hipStream_t s);
hipError_t launch_debug_fp64_math(int fn, const double* x, uint32_t n, double* out, hipStream_t s);
hipError_t launch_debug_draws(const McDesc* desc, uint64_t pair0, uint64_t n_pairs, int fp32, double* za,
                              double* zb, double* jump, uint32_t* n_jumps, hipStream_t s);

}  // namespace ope
