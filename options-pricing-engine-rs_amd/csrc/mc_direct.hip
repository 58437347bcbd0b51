// Direct launches of one request: its descriptor and strikes arrive in the
// kernel arguments, so a call whose launch groups each hold one request
// (a single mc_heston / mc_bates call, or a batch of requests of distinct
// configurations) needs no upload, and the slot reduction runs inside the
// same launch.  Built with -ffp-contract=on from the same tile bodies as
// mc_kernels.hip, so a request computes the same bits here as in a planned
// (staged) launch, and block_sum_slots sums its slots in the same order as
// k_reduce.
#include "mc_device.h"
#include "mc_dispatch.h"
#include "mc_launch.h"

namespace ope {
namespace {

template <class Real, int CODE>
__global__ __launch_bounds__(kBlock, (pinned_waves<Real, CODE>())) void k_tiles_direct(const DirectArgs a,
                                                        double2* __restrict__ slots,
                                                        uint32_t* __restrict__ done,
                                                        double2* __restrict__ out) {
    __shared__ double2 lds[kTileLds];
    __shared__ double2 w[kBlock / 64];
    __shared__ uint32_t last;
    run_code<Real, CODE>(a.d, blockIdx.x, a.strikes, slots, lds);
    // Every thread's slot stores are released at agent scope before the
    // block counts itself (a release increment: the other blocks need no
    // acquire); the block that counts last acquires them all below.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0)
        last = __hip_atomic_fetch_add(done, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT) + 1u == gridDim.x;
    __syncthreads();
    if (!last) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    for (uint32_t k = 0; k < a.d.n_strikes; ++k) {
        const double2 r = block_sum_slots(slots, a.d.slot_off + (uint64_t)k * a.d.n_tiles, a.d.n_tiles, w);
        if (threadIdx.x == 0) out[a.d.out_off + k] = r;
        __syncthreads();  // w is rewritten by the next strike's sum
    }
    // the next launch on this counter follows in stream order
    if (threadIdx.x == 0) __hip_atomic_store(done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class Real, int CODE = 0>
hipError_t launch_direct_code(uint32_t code, const DirectLaunch& L, hipStream_t s) {
    if constexpr (CODE < (int)kNumCodes) {
        if (code == (uint32_t)CODE) {
            hipLaunchKernelGGL((k_tiles_direct<Real, CODE>), dim3(L.args.d.n_tiles), dim3(kBlock), 0, s, L.args,
                               L.slots, L.done, L.out);
            return hipGetLastError();
        }
        return launch_direct_code<Real, CODE + 1>(code, L, s);
    } else {
        return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_direct(const DirectLaunch& L, hipStream_t s) {
    if (L.args.d.n_tiles == 0 || L.args.d.n_strikes > kDirectStrikes) return hipErrorInvalidValue;
    const uint32_t code = config_code(L.args.d);
    return L.fp32 ? launch_direct_code<float>(code, L, s) : launch_direct_code<double>(code, L, s);
}

}  // namespace ope
