// Replay-mode kernels: the same path code as the Philox kernels, fed with
// caller-supplied draws (za, zb, jump_sum per pair-step, the reference's
// StepDraw, mc.rs:126-127).  Built with -ffp-contract=off so that, given the
// reference's draws, the FP64 path repeats the reference's unfused
// arithmetic operation for operation.
#include "mc_device.h"
#include "mc_launch.h"

namespace ope {
namespace {

struct ReplayMaker {
    const double* za;
    const double* zb;
    const double* jump;
    __device__ __forceinline__ ReplaySource operator()(const McDesc& d, uint64_t pair, uint32_t) const {
        ReplaySource s;
        s.za = za;
        s.zb = zb;
        s.jump = jump;
        s.n_pairs = d.n_pairs;
        s.pair = pair;
        return s;
    }
};

// CODE is a config code with ppt == 1 (even codes).
template <class Real, int CODE>
__global__ __launch_bounds__(kBlock) void k_replay(const McDesc* __restrict__ desc, const double* __restrict__ za,
                                                   const double* __restrict__ zb, const double* __restrict__ jump,
                                                   const double* __restrict__ strikes, double2* __restrict__ slots,
                                                   double* __restrict__ pay_out, double* __restrict__ state_out) {
    constexpr int kPayoff = (CODE >> 1) % 3;
    constexpr bool kJumps = (((CODE >> 1) / 3) & 1) != 0;
    constexpr int kScheme = ((CODE >> 1) / 3) >> 1;
    __shared__ double2 lds[kTileLds];
    price_tile<Real, kScheme, kJumps, kPayoff, 1>(*desc, blockIdx.x, strikes, slots, ReplayMaker{za, zb, jump},
                                                  PathSink{pay_out, state_out}, lds);
}

template <class Real, int CODE = 0>
hipError_t launch_replay_code(const ReplayLaunch& L, hipStream_t s) {
    if constexpr (CODE < (int)kNumCodes) {
        if (L.code == (uint32_t)CODE) {
            hipLaunchKernelGGL((k_replay<Real, CODE>), dim3(L.n_tiles), dim3(kBlock), 0, s, L.desc, L.za, L.zb, L.jump,
                               L.strikes, L.slots, L.path_payoff, L.path_state);
            return hipGetLastError();
        }
        return launch_replay_code<Real, CODE + 2>(L, s);
    } else {
        return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_replay(const ReplayLaunch& L, hipStream_t s) {
    if (L.n_tiles == 0) return hipSuccess;
    if (L.code & 1) return hipErrorInvalidValue;  // replay runs one pair per thread
    return L.fp32 ? launch_replay_code<float>(L, s) : launch_replay_code<double>(L, s);
}

}  // namespace ope
