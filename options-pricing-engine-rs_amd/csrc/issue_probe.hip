// Issue-cost probe for the instruction classes of the path kernels: how many
// SIMD cycles one wave64 instruction of each class occupies when the SIMD is
// full (8 waves of 256-thread blocks on every SIMD, 8 independent chains
// per lane).  Standalone executable, not part of the library; it prints one
// line per class and the FP32 rate it implies.
//
//   lib/issue_probe            (built by build.py beside libopemc.so)
//
// Each class runs f() on its chains; some f() need a second instruction to
// keep the chain alive (noted in the name: "sqrt + add" issues one v_sqrt_f32
// and one v_add_f32 per application).  "fma x24 + sqrt+add x8" interleaves
// the two classes to show whether transcendentals overlap other VALU work
// (its cycles are per loop iteration, beside the sum of the parts).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <type_traits>

namespace {

constexpr int kChains = 8, kUnroll = 4, kThreads = 256;

struct Fma { __device__ static float f(float x) { return fmaf(x, 0.999f, 1e-3f); } };
struct Mul { __device__ static float f(float x) { return x * 0.999f; } };
struct Add { __device__ static float f(float x) { return x + 1e-3f; } };
struct SqrtAdd { __device__ static float f(float x) { return __builtin_amdgcn_sqrtf(x) + 1.0f; } };
struct RcpAdd { __device__ static float f(float x) { return __builtin_amdgcn_rcpf(x) + 1.0f; } };
struct LogAdd { __device__ static float f(float x) { return __builtin_amdgcn_logf(x) + 2.0f; } };
struct ExpNeg { __device__ static float f(float x) { return __builtin_amdgcn_exp2f(-x); } };
struct Cos { __device__ static float f(float x) { return __builtin_amdgcn_cosf(x); } };

struct MadU64 { __device__ static uint32_t f(uint32_t x, uint32_t k) { return x * 0xD2511F53u + k; } };
struct Bitop3Shr {
    __device__ static uint32_t f(uint32_t x, uint32_t k) { return __builtin_amdgcn_bitop3_b32(x, k, x >> 1, 0x96); }
};
struct CvtXor { __device__ static uint32_t f(uint32_t x, uint32_t k) { return __float_as_uint((float)x) ^ k; } };

struct AddF64 { __device__ static double f(double x) { return x + 1e-3; } };

template <class Op, class T>
__device__ __forceinline__ T apply(T x, uint32_t k) {
    if constexpr (sizeof(T) == 4 && std::is_same<T, uint32_t>::value)
        return Op::f(x, k);
    else
        return Op::f(x);
}

template <class Op, class T>
__global__ __launch_bounds__(kThreads) void k_chain(T* out, int iters, T seed) {
    T a[kChains];
    uint32_t k = (uint32_t)threadIdx.x * 77u + 5u;
    asm volatile("" : "+v"(k));
#pragma unroll
    for (int i = 0; i < kChains; ++i) a[i] = seed + (T)(threadIdx.x % 7) + (T)i;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int u = 0; u < kUnroll; ++u)
#pragma unroll
            for (int i = 0; i < kChains; ++i) a[i] = apply<Op, T>(a[i], k);
    }
    T s = a[0];
#pragma unroll
    for (int i = 1; i < kChains; ++i) s = s + a[i];
    out[blockIdx.x * kThreads + threadIdx.x] = s;
}

// 24 FMAs on 8 chains and 8 sqrt+add on 4 chains per iteration
__global__ __launch_bounds__(kThreads) void k_mix(float* out, int iters, float seed) {
    float a[8], t[4];
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] = seed + (float)(threadIdx.x % 7) + (float)i;
#pragma unroll
    for (int i = 0; i < 4; ++i) t[i] = 1.5f + (float)(threadIdx.x % 5) + (float)i;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int j = 0; j < 24; ++j) a[j & 7] = fmaf(a[j & 7], 0.999f, 1e-3f);
#pragma unroll
        for (int j = 0; j < 8; ++j) t[j & 3] = __builtin_amdgcn_sqrtf(t[j & 3]) + 1.0f;
    }
    float s = 0.0f;
#pragma unroll
    for (int i = 0; i < 8; ++i) s += a[i];
#pragma unroll
    for (int i = 0; i < 4; ++i) s += t[i];
    out[blockIdx.x * kThreads + threadIdx.x] = s;
}

bool ok(hipError_t e, const char* what) {
    if (e != hipSuccess) std::fprintf(stderr, "issue_probe: %s: %s\n", what, hipGetErrorString(e));
    return e == hipSuccess;
}

// best of 5 timed launches, in ms
template <class Launch>
float best_ms(Launch launch) {
    hipEvent_t e0, e1;
    if (!ok(hipEventCreate(&e0), "event") || !ok(hipEventCreate(&e1), "event")) return -1.0f;
    launch();
    float best = -1.0f;
    if (ok(hipDeviceSynchronize(), "warm-up")) {
        for (int r = 0; r < 5; ++r) {
            float ms = 0.0f;
            if (!ok(hipEventRecord(e0, 0), "record")) break;
            launch();
            if (!ok(hipEventRecord(e1, 0), "record") || !ok(hipEventSynchronize(e1), "sync") ||
                !ok(hipEventElapsedTime(&ms, e0, e1), "elapsed"))
                break;
            if (best < 0.0f || ms < best) best = ms;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

}  // namespace

int main() {
    hipDeviceProp_t p;
    if (!ok(hipGetDeviceProperties(&p, 0), "device")) return 1;
    const int cus = p.multiProcessorCount, simds = cus * 4;
    const double ghz = p.clockRate * 1e-6;
    const int blocks = cus * 8;   // 8 blocks of 4 waves per CU: 8 waves per SIMD
    const int iters = 4000;
    void* d = nullptr;
    if (!ok(hipMalloc(&d, (size_t)blocks * kThreads * sizeof(double)), "alloc")) return 1;
    std::printf("# %s, %d CUs, %.2f GHz max; cycles per wave64 instruction group per SIMD at that clock\n",
                p.gcnArchName, cus, ghz);
    // applications of f per SIMD: 8 waves x iters x chains x unroll
    const double apps = 8.0 * iters * kChains * kUnroll;
    auto report = [&](const char* name, float ms, double groups) {
        if (ms < 0.0f) {
            std::printf("%-22s failed\n", name);
            return;
        }
        std::printf("%-22s %8.3f ms %7.2f cycles\n", name, ms, ms * 1e-3 * ghz * 1e9 / groups);
    };
#define CHAIN(Op, T, seed, name)                                                                              \
    report(name,                                                                                               \
           best_ms([&] {                                                                                       \
               hipLaunchKernelGGL((k_chain<Op, T>), dim3(blocks), dim3(kThreads), 0, 0, (T*)d, iters, (T)(seed)); \
           }),                                                                                                 \
           apps)
    CHAIN(Fma, float, 1.25f, "v_fma_f32");
    CHAIN(Mul, float, 1.25f, "v_mul_f32");
    CHAIN(Add, float, 1.25f, "v_add_f32");
    CHAIN(SqrtAdd, float, 1.25f, "v_sqrt_f32 + add");
    CHAIN(RcpAdd, float, 1.25f, "v_rcp_f32 + add");
    CHAIN(LogAdd, float, 1.25f, "v_log_f32 + add");
    CHAIN(ExpNeg, float, 1.25f, "v_exp_f32");
    CHAIN(Cos, float, 0.25f, "v_cos_f32");
    CHAIN(MadU64, uint32_t, 12345u, "v_mad_u64_u32");
    CHAIN(Bitop3Shr, uint32_t, 12345u, "v_bitop3_b32 + shr");
    CHAIN(CvtXor, uint32_t, 12345u, "v_cvt_f32_u32 + xor");
    CHAIN(AddF64, double, 1.25, "v_add_f64");
#undef CHAIN
    const float fma_ms = best_ms([&] {
        hipLaunchKernelGGL((k_chain<Fma, float>), dim3(blocks), dim3(kThreads), 0, 0, (float*)d, iters, 1.25f);
    });
    const double fma_groups = apps;
    const double tflops = 2.0 * 64.0 * fma_groups * simds / (fma_ms * 1e-3) * 1e-12;
    std::printf("v_fma_f32 rate: %.1f TFLOP/s (FP32, no packed math)\n", tflops);
    // 24 FMAs + 8 sqrt + 8 adds per iteration against the parts alone
    const float mix = best_ms([&] {
        hipLaunchKernelGGL(k_mix, dim3(blocks), dim3(kThreads), 0, 0, (float*)d, iters, 1.25f);
    });
    report("fma x24 + sqrt+add x8", mix, 8.0 * iters);  // cycles per iteration: compare with the parts'
    const float sq = best_ms([&] {
        hipLaunchKernelGGL((k_chain<SqrtAdd, float>), dim3(blocks), dim3(kThreads), 0, 0, (float*)d, iters, 1.25f);
    });
    if (fma_ms > 0.0f && sq > 0.0f)
        std::printf("%-22s %8s    %7.2f cycles if the classes do not overlap\n", "  sum of the parts", "",
                    (24.0 * fma_ms + 8.0 * sq) * 1e-3 * ghz * 1e9 / apps);
    (void)hipFree(d);
    std::printf("done\n");
    return 0;
}
