// Philox-mode pricing kernels, the slot reduction and the generator probes.
// Built with -ffp-contract=on: a*b+c is fused only inside one source
// expression, so a request's arithmetic is the same in the specialised and
// in the mixed-batch kernel, and results do not depend on what shares a
// launch.
#include "mc_device.h"
#include "mc_dispatch.h"
#include "mc_launch.h"

namespace ope {
namespace {

// One block per tile; every request in the launch has config code CODE.
template <class Real, int CODE>
__global__ __launch_bounds__(kBlock, (pinned_waves<Real, CODE>())) void k_tiles_fixed(const McDesc* __restrict__ descs,
                                                        const uint2* __restrict__ tiles,
                                                        const double* __restrict__ strikes,
                                                        double2* __restrict__ slots) {
    __shared__ double2 lds[kTileLds];
    const uint2 t = tiles[blockIdx.x];
    run_code<Real, CODE>(descs[t.x], t.y, strikes, slots, lds);
}

// One block per tile; each tile's request picks its body (uniform branch).
template <class Real>
__global__ __launch_bounds__(kBlock, (std::is_same<Real, float>::value ? kPinnedWavesMixedF32 : 1)) void k_tiles_mixed(const McDesc* __restrict__ descs,
                                                        const uint2* __restrict__ tiles,
                                                        const double* __restrict__ strikes,
                                                        double2* __restrict__ slots) {
    __shared__ double2 lds[kTileLds];
    const uint2 t = tiles[blockIdx.x];
    const McDesc& d = descs[t.x];
    dispatch<Real, 0>(config_code(d), d, t.y, strikes, slots, lds);
}

// Sum one (request, strike) over its tiles in tile order: thread i takes
// tiles i, i+256, ... then the fixed butterfly / wave order.
__global__ __launch_bounds__(kBlock) void k_reduce(const McDesc* __restrict__ descs, const uint2* __restrict__ jobs,
                                                   const double2* __restrict__ slots, double2* __restrict__ out) {
    __shared__ double2 w[kBlock / 64];
    const uint2 jb = jobs[blockIdx.x];
    const McDesc& d = descs[jb.x];
    const double2 r = block_sum_slots(slots, d.slot_off + (uint64_t)jb.y * d.n_tiles, d.n_tiles, w);
    if (threadIdx.x == 0) out[d.out_off + jb.y] = r;
}

// Odd counters go through the split arrangement of the xors
// (philox4x32_10<true>), even ones through the other (the QE loops'),
// so the known-answer checks hold both to the same bits.
__global__ void k_debug_philox(uint32_t k0, uint32_t k1, const uint4* __restrict__ ctr, uint4* __restrict__ out,
                               uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (i & 1u) ? philox4x32_10<true>(ctr[i], k0, k1) : philox4x32_10<false>(ctr[i], k0, k1);
}

// The draws of the pricing loops (philox_steps, as sim_pair_f64 / sim_pair_f32
// consume them), one thread per pair, written [step][pair]; FP32 jumps are
// converted back from log2 to ln units.
__global__ void k_debug_draws(const McDesc* __restrict__ desc, uint64_t pair0, uint64_t n_pairs, int fp32,
                              double* __restrict__ za, double* __restrict__ zb, double* __restrict__ jump,
                              uint32_t* __restrict__ nj) {
    const McDesc& d = *desc;
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_pairs) return;
    const PairStream st = make_stream(d, pair0 + i);
    uint32_t s = 0;
    auto rec = [&](auto a, auto b, auto j, uint32_t n, uint32_t) {
        const uint64_t o = (uint64_t)s * n_pairs + i;
        za[o] = (double)a;
        zb[o] = (double)b;
        jump[o] = fp32 ? (double)j * kLn2 : (double)j;
        if (nj) nj[o] = n;
        ++s;
    };
    if (fp32) {
        if (d.jumps)
            philox_steps<true, true>(d, st, rec);
        else
            philox_steps<true, false>(d, st, rec);
    } else {
        if (d.jumps)
            philox_steps<false, true>(d, st, rec);
        else
            philox_steps<false, false>(d, st, rec);
    }
}

// The FP64 Philox path's elementary functions (fp64_math.h) on given inputs:
// fn 0 log_f64, 1 exp_f64, 2 / 3 sine / cosine of sincos_rev23 (x holds the
// 23-bit integer angle), 4 rcp_f64, 5 rsqrt_f64, 6 sqrt_f64, 7
// ncdf_neg_abs_f64.
__global__ void k_debug_fp64_math(int fn, const double* __restrict__ x, uint32_t n, double* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = x[i];
    double r;
    if (fn == 0) {
        r = log_f64(v);
    } else if (fn == 1) {
        r = exp_f64(v);
    } else if (fn == 2 || fn == 3) {
        double sn, cs;
        sincos_rev23((uint32_t)v, sn, cs);
        r = fn == 2 ? sn : cs;
    } else if (fn == 4) {
        r = rcp_f64(v);
    } else if (fn == 5) {
        r = rsqrt_f64(v);
    } else if (fn == 6) {
        r = sqrt_f64(v);
    } else {
        r = ncdf_neg_abs_f64(v);
    }
    out[i] = r;
}

template <class Real, int CODE = 0>
hipError_t launch_fixed(uint32_t code, const TileLaunch& L, hipStream_t s) {
    if constexpr (CODE < (int)kNumCodes) {
        if (code == (uint32_t)CODE) {
            hipLaunchKernelGGL((k_tiles_fixed<Real, CODE>), dim3(L.n_tiles), dim3(kBlock), 0, s, L.descs, L.tiles,
                               L.strikes, L.slots);
            return hipGetLastError();
        }
        return launch_fixed<Real, CODE + 1>(code, L, s);
    } else {
        return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_tiles(const TileLaunch& L, hipStream_t s) {
    if (L.n_tiles == 0) return hipSuccess;
    if (L.fixed) return L.fp32 ? launch_fixed<float>(L.code, L, s) : launch_fixed<double>(L.code, L, s);
    if (L.fp32)
        hipLaunchKernelGGL((k_tiles_mixed<float>), dim3(L.n_tiles), dim3(kBlock), 0, s, L.descs, L.tiles, L.strikes,
                           L.slots);
    else
        hipLaunchKernelGGL((k_tiles_mixed<double>), dim3(L.n_tiles), dim3(kBlock), 0, s, L.descs, L.tiles, L.strikes,
                           L.slots);
    return hipGetLastError();
}

hipError_t launch_reduce(const McDesc* descs, const uint2* jobs, uint32_t n_jobs, const double2* slots, double2* out,
                         hipStream_t s) {
    if (n_jobs == 0) return hipSuccess;
    hipLaunchKernelGGL(k_reduce, dim3(n_jobs), dim3(kBlock), 0, s, descs, jobs, slots, out);
    return hipGetLastError();
}

hipError_t launch_debug_philox(uint64_t seed, const uint4* ctr, uint4* out, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_debug_philox, dim3((n + 255) / 256), dim3(256), 0, s, (uint32_t)seed,
                       (uint32_t)(seed >> 32), ctr, out, n);
    return hipGetLastError();
}

hipError_t launch_debug_fp64_math(int fn, const double* x, uint32_t n, double* out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_debug_fp64_math, dim3((n + 255) / 256), dim3(256), 0, s, fn, x, n, out);
    return hipGetLastError();
}

hipError_t launch_debug_draws(const McDesc* desc, uint64_t pair0, uint64_t n_pairs, int fp32, double* za, double* zb,
                              double* jump, uint32_t* n_jumps, hipStream_t s) {
    if (n_pairs == 0) return hipSuccess;
    hipLaunchKernelGGL(k_debug_draws, dim3((unsigned)((n_pairs + 255) / 256)), dim3(256), 0, s, desc, pair0, n_pairs,
                       fp32, za, zb, jump, n_jumps);
    return hipGetLastError();
}

}  // namespace ope
