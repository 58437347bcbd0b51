// Philox-mode tile bodies and their per-request dispatch, shared by the
// batch kernels (mc_kernels.hip) and the persistent serving kernel
// (mc_server.hip).  Both units are built with -ffp-contract=on, so a request
// computes the same bits in either.
#pragma once

#include "mc_device.h"
#include "mc_launch.h"

namespace ope {
namespace {

struct PhiloxMaker {
    __device__ __forceinline__ PhiloxSource operator()(const McDesc& d, uint64_t pair, uint32_t pair_hi) const {
        PhiloxSource s;
        s.st = make_stream(d, pair);
        s.st.phi = pair_hi;  // equal to pair >> 32, known uniform
        return s;
    }
};

template <int CODE>
struct Code {
    static constexpr int kPPT = (CODE & 1) ? 4 : 1;
    static constexpr int kPayoff = (CODE >> 1) % 3;
    static constexpr bool kJumps = (((CODE >> 1) / 3) & 1) != 0;
    static constexpr int kScheme = ((CODE >> 1) / 3) >> 1;
};

// Waves per SIMD pinned with __launch_bounds__ for the FP32-path up-and-out
// bodies: their cold FP64 re-simulation near the barrier must not raise the
// register budget of the FP32 loops, so each keeps the occupancy it has
// without that code (the rare FP64 pass spills instead).  FP64: the QE
// bodies at 3 waves (unpinned the compiler spends 169-177 VGPRs where 129-163
// suffice without spilling, and the QE pair runs 3.63 ms at 3 waves against
// 4.04 ms at 2) and the Bates Euler bodies at 5 (16.70 against 16.80 ms for
// the surface unpinned; at 6 they fit 80 VGPRs by spilling 28-80 bytes per
// lane and run 16.1 against 15.8 ms for the surface, 70.0 against 60.4 ms
// for the uneven batch).  1: no constraint.
template <class Real, int CODE>
constexpr int pinned_waves() {
    if (!std::is_same<Real, float>::value) {
        if (CODE == 18) return 4;
        if (CODE >= 12) return 3;
        return (CODE == 6 || CODE == 8 || CODE == 10) ? 5 : 1;
    }
    switch (CODE) {
    case 4: return 8;
    case 5: return 4;
    case 10: return 6;
    case 11: return 3;
    case 16: return 8;
    case 17: return 5;
    case 22: return 8;
    case 23: return 4;
    default: return 1;
    }
}
// The mixed FP32 kernel and the FP32 serving kernel, which hold every body.
constexpr int kPinnedWavesMixedF32 = 3;
constexpr int kPinnedWavesServerF32 = 4;

template <class Real, int CODE>
__device__ __forceinline__ void run_code(const McDesc& d, uint32_t tile, const double* strikes, double2* slots,
                                         double2* lds) {
    using C = Code<CODE>;
    price_tile<Real, C::kScheme, C::kJumps, C::kPayoff, C::kPPT>(d, tile, strikes, slots, PhiloxMaker{},
                                                                 PathSink{nullptr, nullptr}, lds);
}

// Runs the body of config code `code` (uniform).  STEP = 2 from CODE = 0
// instantiates only the one-pair-per-thread bodies.
template <class Real, int CODE, int STEP = 1>
__device__ __forceinline__ void dispatch(uint32_t code, const McDesc& d, uint32_t tile, const double* strikes,
                                         double2* slots, double2* lds) {
    if constexpr (CODE < (int)kNumCodes) {
        if (code == (uint32_t)CODE) {
            run_code<Real, CODE>(d, tile, strikes, slots, lds);
            return;
        }
        dispatch<Real, CODE + STEP, STEP>(code, d, tile, strikes, slots, lds);
    }
}

}  // namespace
}  // namespace ope
