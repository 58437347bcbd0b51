// C ABI of the Monte Carlo engine (include/ope_mc.h): validation, request
// descriptors, tile planning, launches and the final mean / standard error.
//
// mc_price (mc.rs:148-207) is split here into:
//   make_desc     the loop invariants of mc.rs:153-157 and run_path's QE
//                 coefficients (mc.rs:230-238), computed once in FP64
//   plan          pairs -> tiles of kBlock*ppt pairs -> FP64 partial slots
//   run_plan      direct launches (one request per launch group, descriptor
//                 and strikes in the kernel arguments, the reduction inside
//                 the launch), or staged launches + one slot reduction launch
//   finish        mean and std_error exactly as mc.rs:198-204
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ope_mc.h"
#include "engine.h"
#include "mc_desc.h"
#include "mc_launch.h"

namespace ope {
namespace host {

constexpr double kLog2eD = 1.4426950408889634;
// Pairs at which a request uses 4 pairs per thread.  Below it one pair per
// thread gives 4x the tiles, so the last round of tiles on the GPU is
// fuller (the 8 x 2^21-pair Bates surface: 4.99 -> 4.94 ms with its strike
// loop's cost per tile); above it the tile count is large either way.
constexpr uint64_t kPPT4Threshold = 1ull << 22;
constexpr uint64_t kMaxPairs = 1ull << 40;
constexpr uint64_t kMaxSteps = 1ull << 24;
constexpr double kMaxPoissonMean = 64.0;

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

// Device buffer that only grows.
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t n) {
        if (n <= cap) return hipSuccess;
        if (p) {
            hipError_t e = hipFree(p);
            if (e != hipSuccess) return e;
            p = nullptr;
            cap = 0;
        }
        size_t want = std::max<size_t>(n, 256);
        hipError_t e = hipMalloc(&p, want);
        if (e == hipSuccess) cap = want;
        return e;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// Pinned host buffer that only grows: one call's inputs are staged here and
// cross PCIe in one copy, and its results come back through it.
struct HostBuf {
    void* p = nullptr;
    size_t cap = 0;
    unsigned flags = hipHostMallocDefault;
    hipError_t ensure(size_t n) {
        if (n <= cap) return hipSuccess;
        if (p) {
            hipError_t e = hipHostFree(p);
            if (e != hipSuccess) return e;
            p = nullptr;
            cap = 0;
        }
        size_t want = std::max<size_t>(n, 4096);
        hipError_t e = hipHostMalloc(&p, want, flags);
        if (e == hipSuccess) cap = want;
        return e;
    }
};

// Launch groups of one call beyond the first run on these streams, so that
// a group's last tiles overlap the next group's first ones instead of
// leaving CUs idle between two kernels (the stream, plus these, stays
// within the 4 hardware queues a process gets by default).
constexpr int kAuxStreams = 3;

struct Context {
    std::mutex mu;
    bool tried = false;
    int init_status = OPE_OK;
    hipStream_t stream = nullptr;
    hipStream_t aux[kAuxStreams] = {};
    hipEvent_t join[1 + kAuxStreams] = {};
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evk = nullptr;
    DevBuf descs, tiles, strikes, slots, jobs, out;
    DevBuf stage;           // a planned call's descriptors, tiles, strikes and jobs
    HostBuf h_stage;
    // the reduction kernel's results, written by the device straight into
    // host memory (coherent, so a synchronised stream has made them visible):
    // no device-to-host copy per call
    HostBuf h_out{nullptr, 0, hipHostMallocCoherent};
    DevBuf rz_a, rz_b, rz_j, path_pay, path_state;
    // completion counters of the direct launches, one per stream: zero
    // between calls (each launch's last block clears its own)
    DevBuf done;
    DevBuf dbg_in, dbg_out, dbg_a, dbg_b, dbg_j, dbg_n;
};

Context g_ctx;

int ensure_init(Context& c) {
    if (c.tried) return c.init_status;
    c.tried = true;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n < 1) {
        c.init_status = fail(OPE_ERR_NO_DEVICE, "no HIP device available (%s)", hipGetErrorString(e));
        return c.init_status;
    }
    if ((e = hipSetDevice(0)) != hipSuccess || (e = hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreate(&c.ev0)) != hipSuccess || (e = hipEventCreate(&c.ev1)) != hipSuccess ||
        (e = hipEventCreate(&c.evk)) != hipSuccess) {
        c.init_status = fail(OPE_ERR_NO_DEVICE, "HIP initialisation failed: %s", hipGetErrorString(e));
        return c.init_status;
    }
    for (int i = 0; i <= kAuxStreams; ++i) {
        if ((i < kAuxStreams && (e = hipStreamCreateWithFlags(&c.aux[i], hipStreamNonBlocking)) != hipSuccess) ||
            (e = hipEventCreateWithFlags(&c.join[i], hipEventDisableTiming)) != hipSuccess) {
            c.init_status = fail(OPE_ERR_NO_DEVICE, "HIP initialisation failed: %s", hipGetErrorString(e));
            return c.init_status;
        }
    }
    if ((e = c.done.ensure((1 + kAuxStreams) * sizeof(uint32_t))) != hipSuccess ||
        (e = hipMemsetAsync(c.done.p, 0, (1 + kAuxStreams) * sizeof(uint32_t), c.stream)) != hipSuccess ||
        (e = hipStreamSynchronize(c.stream)) != hipSuccess) {
        c.init_status = fail(OPE_ERR_NO_DEVICE, "HIP initialisation failed: %s", hipGetErrorString(e));
        return c.init_status;
    }
    c.init_status = OPE_OK;
    return OPE_OK;
}

bool finite_all(std::initializer_list<double> xs) {
    for (double x : xs)
        if (!std::isfinite(x)) return false;
    return true;
}

// Validates a request and fills its descriptor (everything except the
// planning fields).  Expressions follow the reference's evaluation order.
int make_desc(const Request& r, McDesc& d) {
    std::memset(&d, 0, sizeof d);
    const ope_heston_params& hp = r.hp;
    if (!finite_all({r.spot, r.expiry, r.rate, r.div_yield, hp.v0, hp.kappa, hp.theta, hp.sigma, hp.rho,
                     r.payoff.strike}))
        return fail(OPE_ERR_INVALID_ARGUMENT, "non-finite argument");
    if (!(r.spot > 0.0)) return fail(OPE_ERR_INVALID_ARGUMENT, "spot must be > 0 (got %g)", r.spot);
    if (!(r.expiry >= 0.0)) return fail(OPE_ERR_INVALID_ARGUMENT, "expiry must be >= 0 (got %g)", r.expiry);
    if (r.cfg.n_steps < 1 || r.cfg.n_steps > kMaxSteps)
        return fail(OPE_ERR_INVALID_ARGUMENT, "n_steps must be in [1, %llu]", (unsigned long long)kMaxSteps);
    if (r.cfg.scheme != OPE_SCHEME_FULL_TRUNCATION_EULER && r.cfg.scheme != OPE_SCHEME_QUADRATIC_EXPONENTIAL)
        return fail(OPE_ERR_INVALID_ARGUMENT, "unknown variance scheme %d", r.cfg.scheme);
    if (r.payoff.kind < OPE_PAYOFF_EUROPEAN || r.payoff.kind > OPE_PAYOFF_UP_AND_OUT)
        return fail(OPE_ERR_INVALID_ARGUMENT, "unknown payoff kind %d", r.payoff.kind);
    if (r.payoff.opt_type != OPE_CALL && r.payoff.opt_type != OPE_PUT)
        return fail(OPE_ERR_INVALID_ARGUMENT, "unknown option type %d", r.payoff.opt_type);
    if (r.payoff.kind == OPE_PAYOFF_UP_AND_OUT && !finite_all({r.payoff.barrier, r.payoff.rebate}))
        return fail(OPE_ERR_INVALID_ARGUMENT, "non-finite barrier or rebate");
    // pairs = floor(n_paths / 2) with antithetic pairs: an odd path is dropped (mc.rs:160-161)
    const uint64_t per_pair = r.cfg.antithetic ? 2 : 1;
    const uint64_t n_pairs = r.cfg.n_paths / per_pair;
    if (n_pairs < 1)
        return fail(OPE_ERR_INVALID_ARGUMENT, "n_paths=%llu gives no path%s", (unsigned long long)r.cfg.n_paths,
                    r.cfg.antithetic ? " pair (antithetic needs n_paths >= 2)" : "");
    if (n_pairs > kMaxPairs) return fail(OPE_ERR_INVALID_ARGUMENT, "n_paths too large");
    const bool qe = r.cfg.scheme == OPE_SCHEME_QUADRATIC_EXPONENTIAL;
    // theta < 0 would make the conditional mean m negative, which the QE
    // steps of the Philox kernels (v >= 0 throughout, m w = sqrt(m^2 - s2/2))
    // do not represent
    if (qe && !(hp.sigma > 0.0 && hp.kappa > 0.0 && hp.theta >= 0.0))
        return fail(OPE_ERR_INVALID_ARGUMENT, "QE scheme needs sigma > 0, kappa > 0 and theta >= 0");
    if (r.has_jump_params) {
        if (!finite_all({r.lambda, r.mu_j, r.sigma_j}))
            return fail(OPE_ERR_INVALID_ARGUMENT, "non-finite jump parameters");
        if (!(r.lambda >= 0.0)) return fail(OPE_ERR_INVALID_ARGUMENT, "lambda must be >= 0");
    }

    // mc.rs:153-157
    const double dt = r.expiry / (double)r.cfg.n_steps;
    const double sqdt = std::sqrt(dt);
    const double k_bar = r.has_jump_params ? std::exp(r.mu_j + 0.5 * r.sigma_j * r.sigma_j) - 1.0 : 0.0;
    const double lambda = r.has_jump_params ? r.lambda : 0.0;
    const double drift = r.rate - r.div_yield - lambda * k_bar;
    const bool jumps = lambda > 0.0;
    const double pois_mu = lambda * dt;
    if (jumps && !(pois_mu <= kMaxPoissonMean))
        return fail(OPE_ERR_INVALID_ARGUMENT, "lambda*dt = %g exceeds the Poisson sampler limit %g", pois_mu,
                    kMaxPoissonMean);

    d.seed = r.cfg.seed;
    d.n_pairs = n_pairs;
    d.n_steps = (uint32_t)r.cfg.n_steps;
    d.scheme = qe ? kSchemeQE : kSchemeEuler;
    d.payoff = (uint32_t)r.payoff.kind;
    d.opt_put = r.payoff.opt_type == OPE_PUT;
    d.antithetic = r.cfg.antithetic ? 1 : 0;
    d.jumps = jumps ? 1 : 0;

    d.spot = r.spot;
    d.ln_spot = std::log(r.spot);
    d.disc = std::exp(-r.rate * r.expiry);
    d.dt = dt;
    d.sqdt = sqdt;
    d.drift = drift;
    d.v0 = hp.v0;
    d.kappa = hp.kappa;
    d.theta = hp.theta;
    d.sigma = hp.sigma;
    d.rho = hp.rho;
    d.sqrt_1m_rho2 = std::sqrt(1.0 - hp.rho * hp.rho);
    d.barrier = r.payoff.kind == OPE_PAYOFF_UP_AND_OUT ? r.payoff.barrier : INFINITY;
    d.rebate = r.payoff.kind == OPE_PAYOFF_UP_AND_OUT ? r.payoff.rebate : 0.0;
    d.lb2_barrier = std::log2(d.barrier / r.spot);

    if (qe) {  // mc.rs:230-238 and 291-294
        const double rho = hp.rho, g1 = 0.5, g2 = 0.5;
        const double k0 = -rho * hp.kappa * hp.theta * dt / hp.sigma;
        d.k1 = g1 * dt * (hp.kappa * rho / hp.sigma - 0.5) - rho / hp.sigma;
        d.k2 = g2 * dt * (hp.kappa * rho / hp.sigma - 0.5) + rho / hp.sigma;
        d.k3 = g1 * dt * (1.0 - rho * rho);
        d.k4 = g2 * dt * (1.0 - rho * rho);
        d.qe_a0 = drift * dt + k0;
        d.ekt = std::exp(-hp.kappa * dt);
        d.one_m_ekt = 1.0 - d.ekt;
        d.qe_c2 = (hp.theta * hp.sigma * hp.sigma / (2.0 * hp.kappa)) * (d.one_m_ekt * d.one_m_ekt);
        // FP32-path folding: m = m1 v + m0, s2 = s1 v + s0, log2-scaled price terms
        d.qe_m1 = (float)d.ekt;
        d.qe_m0 = (float)(hp.theta * d.one_m_ekt);
        d.qe_s1 = (float)(0.5 * (hp.sigma * hp.sigma * d.ekt * d.one_m_ekt / hp.kappa));
        d.qe_s0 = (float)(0.5 * d.qe_c2);
        d.qe_q0 = (float)(kLog2eD * d.qe_a0);
        d.qe_q1 = (float)(kLog2eD * d.k1);
        d.qe_q2 = (float)(kLog2eD * d.k2);
        // k3 == k4 (gamma1 == gamma2), so the FP32 path's variance term is
        // sqrt(max(k3, 0)) sqrt(v + v'): the clamp at 0 only ever acts for k3 < 0
        d.qe_sq3 = (float)(kLog2eD * std::sqrt(std::fmax(d.k3, 0.0)));
        d.qd_m0 = hp.theta * d.one_m_ekt;
        d.qd_s1 = 0.5 * (hp.sigma * hp.sigma * d.ekt * d.one_m_ekt / hp.kappa);
        d.qd_s0 = 0.5 * d.qe_c2;
        d.qd_sq3 = std::sqrt(std::fmax(d.k3, 0.0));
        // regrouping threshold: u(v) = (s1 v + s0) / (ekt v + m0)^2 = c, c = 0.1,
        // i.e. c ekt^2 v^2 + (2 c ekt m0 - s1) v + c m0^2 - s0 = 0; u falls
        // with v, so v below the positive root has u > c (none if u(0) <= c)
        const double c = 0.1, qa = c * d.ekt * d.ekt, qb = 2.0 * c * d.ekt * d.qd_m0 - d.qd_s1,
                     qc = c * d.qd_m0 * d.qd_m0 - d.qd_s0;
        d.qe_vlow = qc < 0.0 ? (float)((-qb + std::sqrt(qb * qb - 4.0 * qa * qc)) / (2.0 * qa)) : 0.0f;
    }
    d.ed_a0 = drift * dt;
    d.ed_a1 = -0.5 * dt;
    d.ed_b0 = hp.kappa * hp.theta * dt;
    d.ed_c1 = 1.0 - hp.kappa * dt;
    d.ed_sb = hp.sigma * sqdt;
    d.f_v0 = (float)hp.v0;
    d.eu_a0 = (float)(kLog2eD * drift * dt);
    d.eu_a1 = (float)(-0.5 * kLog2eD * dt);
    d.eu_a2 = (float)(kLog2eD * sqdt);
    d.eu_b0 = (float)(hp.kappa * hp.theta * dt);
    d.eu_b1 = (float)(-hp.kappa * dt);
    d.eu_c1 = (float)(1.0 - hp.kappa * dt);
    {
        const double a2 = kLog2eD * sqdt;
        d.eu_r2 = (float)(-2.0 * std::log(2.0) * a2 * a2);
        d.eu_kb = (float)(hp.sigma / kLog2eD);  // b2 / a2
        // |rho| > 1 gives NaN here as sqrt(1 - rho^2) does in the reference
        d.eu_phi = (float)(std::acos(hp.rho) / 6.283185307179586);
    }
    d.eu_b2 = (float)(hp.sigma * sqdt);
    d.f_rho = (float)hp.rho;
    d.f_rhoc = (float)d.sqrt_1m_rho2;

    if (jumps) {
        d.pois_mu = pois_mu;
        d.pois_e0 = std::exp(-pois_mu);
        const double thr = std::ceil(d.pois_e0 * 4294967296.0 - 0.5);
        d.pois_thr0 = thr >= 4294967295.0 ? 0xFFFFFFFFu : (thr <= 0.0 ? 0u : (uint32_t)thr);
        d.pois_gate = d.pois_thr0 >> 23;
        // F1 = e0 + e0 mu / 1, the CDF after the device walk's first step,
        // with that step's roundings (the division by 1 is exact); u < F1
        // <=> w < ceil(F1 2^32 - 1/2), both sides exact.  Clamped below
        // 2^32: the largest w then takes the walk, which gives the same count.
        const double f1 = [&] {
#pragma clang fp contract(off)
            const double p1 = d.pois_e0 * pois_mu;
            return d.pois_e0 + p1;
        }();
        const double thr1 = std::ceil(f1 * 4294967296.0 - 0.5);
        d.pois_thr1 = thr1 >= 4294967295.0 ? 0xFFFFFFFFu : (thr1 <= 0.0 ? 0u : (uint32_t)thr1);
        d.mu_j = r.mu_j;
        d.sigma_j = r.sigma_j;
        d.f_mu_j2 = (float)(kLog2eD * r.mu_j);
        d.f_sigma_j2 = (float)(kLog2eD * r.sigma_j);
    }
    return OPE_OK;
}

// One pricing job: a descriptor plus the strikes it prices.
struct Job {
    McDesc d;
    std::vector<double> strikes;
};

struct Plan {
    std::vector<McDesc> descs;
    std::vector<uint2> tiles;
    std::vector<double> strikes;
    std::vector<uint2> jobs;  // (request, strike)
    uint64_t n_slots = 0;
    uint32_t n_out = 0;
    // Launch groups over `tiles`: each group is one kernel launch; code ==
    // kNumCodes marks the mixed kernel; req is the group's request when it
    // has only one (else ~0u).
    struct Group {
        uint32_t code, first, count, req;
    };
    std::vector<Group> groups;
};

// Up to this many distinct request configurations run as one specialised
// launch each; more than that share one mixed launch.
constexpr size_t kMaxFixedGroups = 6;

// Pairs per thread and tile count of a request (a function of the request
// alone, so a request computes the same bits in every entry point).
int set_tiling(McDesc& d) {
    d.ppt = d.n_pairs >= kPPT4Threshold ? 4 : 1;
    const uint64_t per_tile = (uint64_t)kBlock * d.ppt;
    const uint64_t nt = (d.n_pairs + per_tile - 1) / per_tile;
    if (nt > 0x7fffffffull) return fail(OPE_ERR_INVALID_ARGUMENT, "too many tiles");
    d.n_tiles = (uint32_t)nt;
    return OPE_OK;
}

int set_single_tiling(McDesc& d) {
    const int rc = set_tiling(d);
    if (rc) return rc;
    d.n_strikes = 1;
    d.strike_off = 0;
    d.out_off = 0;
    d.slot_off = 0;
    return OPE_OK;
}

int plan(std::vector<Job>& in, Plan& p) {
    p = Plan();
    uint64_t slot = 0;
    uint32_t out = 0;
    for (size_t i = 0; i < in.size(); ++i) {
        McDesc& d = in[i].d;
        const size_t ns = in[i].strikes.size();
        if (ns < 1 || ns > (size_t)kMaxStrikes)
            return fail(OPE_ERR_INVALID_ARGUMENT, "strike count %zu outside [1, %d]", ns, kMaxStrikes);
        for (double k : in[i].strikes)
            if (!std::isfinite(k)) return fail(OPE_ERR_INVALID_ARGUMENT, "non-finite strike");
        int rc = set_tiling(d);
        if (rc) return rc;
        const uint64_t nt = d.n_tiles;
        d.n_strikes = (uint32_t)ns;
        d.strike_off = (uint32_t)p.strikes.size();
        d.out_off = out;
        d.slot_off = slot;
        slot += nt * ns;
        out += (uint32_t)ns;
        p.strikes.insert(p.strikes.end(), in[i].strikes.begin(), in[i].strikes.end());
        p.descs.push_back(d);
        for (uint32_t k = 0; k < ns; ++k) p.jobs.push_back(make_uint2((uint32_t)i, k));
    }
    p.n_slots = slot;
    p.n_out = out;
    uint64_t total = 0;
    for (const McDesc& d : p.descs) total += d.n_tiles;
    if (total > 0x7fffffffull) return fail(OPE_ERR_INVALID_ARGUMENT, "too many tiles in one launch");

    // Tiles of the heaviest requests first, so long tiles start early; the
    // order of execution never changes a result.
    auto cost = [&](uint32_t i) {
        const McDesc& d = p.descs[i];
        return (double)d.n_steps * d.ppt * (d.scheme == kSchemeQE ? 1.5 : 1.0) * (d.jumps ? 1.3 : 1.0);
    };
    std::vector<uint32_t> order(in.size());
    for (uint32_t i = 0; i < order.size(); ++i) order[i] = i;
    std::vector<uint32_t> codes;
    for (const McDesc& d : p.descs)
        if (std::find(codes.begin(), codes.end(), config_code(d)) == codes.end()) codes.push_back(config_code(d));
    const bool grouped = codes.size() <= kMaxFixedGroups;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        if (grouped) {
            const uint32_t ca = config_code(p.descs[a]), cb = config_code(p.descs[b]);
            if (ca != cb) return ca < cb;
        }
        return cost(a) > cost(b);
    });
    p.tiles.reserve(total);
    for (uint32_t i : order) {
        const uint32_t code = grouped ? config_code(p.descs[i]) : kNumCodes;
        if (p.groups.empty() || p.groups.back().code != code)
            p.groups.push_back(Plan::Group{code, (uint32_t)p.tiles.size(), 0, i});
        else
            p.groups.back().req = ~0u;
        for (uint32_t t = 0; t < p.descs[i].n_tiles; ++t) p.tiles.push_back(make_uint2(i, t));
        p.groups.back().count += p.descs[i].n_tiles;
    }
    return OPE_OK;
}

template <class T>
hipError_t upload(DevBuf& b, const std::vector<T>& v, hipStream_t s) {
    hipError_t e = b.ensure(v.size() * sizeof(T));
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
}

// Raw (sum, sum_sq) per (request, strike) after the launch.
struct RawOut {
    std::vector<double2> sums;
    double device_ms = 0.0;  // pricing + reduction
    double kernel_ms = 0.0;  // pricing kernel alone
};

// A plan runs as direct launches when each of its groups is one request of
// one configuration with at most kDirectStrikes strikes, one group per
// stream: its descriptor and strikes go in the kernel arguments and the
// launch reduces its own slots, so the call makes no upload and no
// reduction launch.  The arithmetic and the reduction order are those of
// the staged launches below.
bool is_direct(const Plan& p) {
    if (p.groups.empty() || p.groups.size() > (size_t)(1 + kAuxStreams)) return false;
    for (const Plan::Group& g : p.groups)
        if (g.code >= kNumCodes || g.req == ~0u || p.descs[g.req].n_strikes > kDirectStrikes) return false;
    return true;
}

int run_direct(Context& c, const Plan& p, int fp32, RawOut& r) {
    hipStream_t s = c.stream;
    const size_t b_out = p.n_out * sizeof(double2);
    HIP_TRY(c.h_out.ensure(b_out));
    HIP_TRY(c.slots.ensure(p.n_slots * sizeof(double2)));
    HIP_TRY(hipEventRecord(c.ev0, s));
    const int n = (int)p.groups.size();
    auto lane_stream = [&](int lane) { return lane == 0 ? s : c.aux[lane - 1]; };
    for (int gi = 0; gi < n; ++gi) {
        const McDesc& d = p.descs[p.groups[gi].req];
        DirectLaunch L;
        L.args.d = d;
        L.args.d.strike_off = 0;
        for (uint32_t k = 0; k < kDirectStrikes; ++k)
            L.args.strikes[k] = k < d.n_strikes ? p.strikes[d.strike_off + k] : 0.0;
        L.slots = c.slots.as<double2>();
        L.done = c.done.as<uint32_t>() + gi;
        L.out = static_cast<double2*>(c.h_out.p);
        L.fp32 = fp32;
        if (gi > 0) HIP_TRY(hipStreamWaitEvent(c.aux[gi - 1], c.ev0, 0));
        HIP_TRY(launch_direct(L, lane_stream(gi)));
    }
    hipStream_t fin = lane_stream(n - 1);
    for (int i = 0; i < n - 1; ++i) {
        HIP_TRY(hipEventRecord(c.join[i], lane_stream(i)));
        HIP_TRY(hipStreamWaitEvent(fin, c.join[i], 0));
    }
    HIP_TRY(hipEventRecord(c.ev1, fin));
    HIP_TRY(hipStreamSynchronize(fin));
    r.sums.resize(p.n_out);
    std::memcpy(r.sums.data(), c.h_out.p, b_out);
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    r.device_ms = ms;
    r.kernel_ms = ms;  // the path kernels, each with its last block's reduction
    return OPE_OK;
}

int run_plan(Context& c, const Plan& p, int fp32, RawOut& r) {
    if (is_direct(p)) return run_direct(c, p, fp32, r);
    hipStream_t s = c.stream;
    // descriptors, tiles, strikes and jobs: one pinned staging area, one copy
    auto at = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t b_desc = p.descs.size() * sizeof(McDesc), b_tiles = p.tiles.size() * sizeof(uint2);
    const size_t b_str = p.strikes.size() * sizeof(double), b_jobs = p.jobs.size() * sizeof(uint2);
    const size_t o_tiles = at(b_desc), o_str = o_tiles + at(b_tiles), o_jobs = o_str + at(b_str);
    const size_t total = o_jobs + b_jobs, b_out = p.n_out * sizeof(double2);
    HIP_TRY(c.h_stage.ensure(total));
    HIP_TRY(c.stage.ensure(total));
    HIP_TRY(c.h_out.ensure(b_out));
    char* h = static_cast<char*>(c.h_stage.p);
    std::memcpy(h, p.descs.data(), b_desc);
    std::memcpy(h + o_tiles, p.tiles.data(), b_tiles);
    std::memcpy(h + o_str, p.strikes.data(), b_str);
    std::memcpy(h + o_jobs, p.jobs.data(), b_jobs);
    HIP_TRY(hipMemcpyAsync(c.stage.p, h, total, hipMemcpyHostToDevice, s));
    char* dv = static_cast<char*>(c.stage.p);
    const McDesc* d_desc = reinterpret_cast<const McDesc*>(dv);
    const uint2* d_tiles = reinterpret_cast<const uint2*>(dv + o_tiles);
    const double* d_str = reinterpret_cast<const double*>(dv + o_str);
    const uint2* d_jobs = reinterpret_cast<const uint2*>(dv + o_jobs);
    HIP_TRY(c.slots.ensure(p.n_slots * sizeof(double2)));
    HIP_TRY(hipEventRecord(c.ev0, s));
    // group i runs on lane i mod (1 + kAuxStreams) (lane 0 = the stream, the
    // others auxiliary streams that start after the uploads, ev0).  Every
    // tile writes its own slots, so the overlap changes no result.
    const int n_lanes = std::min<int>(1 + kAuxStreams, (int)p.groups.size());
    auto lane_stream = [&](int lane) { return lane == 0 ? s : c.aux[lane - 1]; };
    for (size_t gi = 0; gi < p.groups.size(); ++gi) {
        const Plan::Group& g = p.groups[gi];
        const size_t lane = gi % (size_t)(1 + kAuxStreams);
        // an auxiliary lane's wait for the uploads is issued just before its
        // first launch, so the first group's launch is not queued behind it
        if (lane > 0 && gi < (size_t)(1 + kAuxStreams)) HIP_TRY(hipStreamWaitEvent(c.aux[lane - 1], c.ev0, 0));
        TileLaunch L;
        L.descs = d_desc;
        L.tiles = d_tiles + g.first;
        L.n_tiles = g.count;
        L.strikes = d_str;
        L.slots = c.slots.as<double2>();
        L.fp32 = fp32;
        L.fixed = g.code < kNumCodes ? 1 : 0;
        L.code = g.code;
        HIP_TRY(launch_tiles(L, lane_stream((int)lane)));
    }
    // The reduction runs on the lane of the last group, which usually ends
    // last (each group starts as the previous one's last tiles free CUs), so
    // that its wait for the other lanes is already met when it reaches the
    // head of its queue.  Synchronising that lane covers every lane.
    const int last_lane = p.groups.empty() ? 0 : (int)((p.groups.size() - 1) % (size_t)(1 + kAuxStreams));
    hipStream_t fin = lane_stream(last_lane);
    for (int i = 0; i < n_lanes; ++i) {
        if (i == last_lane) continue;
        HIP_TRY(hipEventRecord(c.join[i], lane_stream(i)));
        HIP_TRY(hipStreamWaitEvent(fin, c.join[i], 0));
    }
    HIP_TRY(hipEventRecord(c.evk, fin));
    HIP_TRY(launch_reduce(d_desc, d_jobs, (uint32_t)p.jobs.size(), c.slots.as<double2>(),
                          static_cast<double2*>(c.h_out.p), fin));
    HIP_TRY(hipEventRecord(c.ev1, fin));
    HIP_TRY(hipStreamSynchronize(fin));
    r.sums.resize(p.n_out);
    std::memcpy(r.sums.data(), c.h_out.p, b_out);
    float ms = 0.0f, kms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    HIP_TRY(hipEventElapsedTime(&kms, c.ev0, c.evk));
    r.device_ms = ms;
    r.kernel_ms = kms;
    return OPE_OK;
}

void report_times(const ope_mc_options* o, const RawOut& raw) {
    if (!o) return;
    if (o->device_ms) *o->device_ms = raw.device_ms;
    if (o->kernel_ms) *o->kernel_ms = raw.kernel_ms;
}

// mc.rs:198-204
void finish(const McDesc& d, double2 sums, ope_mc_result* out, ope_mc_stats* st) {
    const uint64_t count = d.n_pairs * (d.antithetic ? 2 : 1);
    const double n = (double)count;
    const double mean = sums.x / n;
    const double var = std::max(sums.y / n - mean * mean, 0.0);
    out->price = mean;
    out->std_error = std::sqrt(var / n);
    if (st) {
        st->sum = sums.x;
        st->sum_sq = sums.y;
        st->count = count;
        st->path_steps = count * (uint64_t)d.n_steps;
    }
}

int check_options(const ope_mc_options* o) {
    if (!o) return OPE_OK;
    if (o->precision != OPE_PRECISION_FP64 && o->precision != OPE_PRECISION_FP32_PATH)
        return fail(OPE_ERR_INVALID_ARGUMENT, "unknown precision %d", o->precision);
    if (o->rng != OPE_RNG_PHILOX && o->rng != OPE_RNG_REPLAY)
        return fail(OPE_ERR_INVALID_ARGUMENT, "unknown rng mode %d", o->rng);
    if (o->flags & ~OPE_FLAG_BARRIER_FP64) return fail(OPE_ERR_INVALID_ARGUMENT, "unknown flags 0x%x", o->flags);
    return OPE_OK;
}

int price_replay(Context& c, Job& job, const ope_mc_options* o, ope_mc_result* out) {
    const ope_replay_draws* rp = o->replay;
    McDesc& d = job.d;
    if (!rp || !rp->za || !rp->zb) return fail(OPE_ERR_INVALID_ARGUMENT, "replay mode needs replay draws");
    if (rp->n_pairs != d.n_pairs || rp->n_steps != d.n_steps)
        return fail(OPE_ERR_INVALID_ARGUMENT, "replay draws are %llux%llu, the request needs %llu pairs x %u steps",
                    (unsigned long long)rp->n_pairs, (unsigned long long)rp->n_steps,
                    (unsigned long long)d.n_pairs, d.n_steps);
    if (d.jumps && !rp->jump_sum) return fail(OPE_ERR_INVALID_ARGUMENT, "replay with jumps needs jump_sum");
    d.ppt = 1;
    d.n_tiles = (uint32_t)((d.n_pairs + kBlock - 1) / kBlock);
    d.n_strikes = 1;
    d.strike_off = 0;
    d.out_off = 0;
    d.slot_off = 0;
    hipStream_t s = c.stream;
    const size_t n = (size_t)d.n_pairs * d.n_steps;
    HIP_TRY(c.rz_a.ensure(n * 8));
    HIP_TRY(c.rz_b.ensure(n * 8));
    HIP_TRY(hipMemcpyAsync(c.rz_a.p, rp->za, n * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(c.rz_b.p, rp->zb, n * 8, hipMemcpyHostToDevice, s));
    const double* dj = nullptr;
    if (d.jumps) {
        HIP_TRY(c.rz_j.ensure(n * 8));
        HIP_TRY(hipMemcpyAsync(c.rz_j.p, rp->jump_sum, n * 8, hipMemcpyHostToDevice, s));
        dj = c.rz_j.as<double>();
    }
    std::vector<McDesc> descs{d};
    std::vector<double> strikes{job.strikes[0]};
    std::vector<uint2> jobs{make_uint2(0, 0)};
    HIP_TRY(upload(c.descs, descs, s));
    HIP_TRY(upload(c.strikes, strikes, s));
    HIP_TRY(upload(c.jobs, jobs, s));
    HIP_TRY(c.slots.ensure((size_t)d.n_tiles * sizeof(double2)));
    HIP_TRY(c.out.ensure(sizeof(double2)));
    double* dpay = nullptr;
    double* dstate = nullptr;
    if (o->path_payoffs) {
        HIP_TRY(c.path_pay.ensure((size_t)d.n_pairs * 2 * 8));
        dpay = c.path_pay.as<double>();
        HIP_TRY(hipMemsetAsync(dpay, 0xff, (size_t)d.n_pairs * 2 * 8, s));
    }
    if (o->path_states) {
        HIP_TRY(c.path_state.ensure((size_t)d.n_pairs * 8 * 8));
        dstate = c.path_state.as<double>();
        HIP_TRY(hipMemsetAsync(dstate, 0xff, (size_t)d.n_pairs * 8 * 8, s));
    }
    ReplayLaunch L;
    L.desc = c.descs.as<McDesc>();
    L.n_tiles = d.n_tiles;
    L.za = c.rz_a.as<double>();
    L.zb = c.rz_b.as<double>();
    L.jump = dj;
    L.strikes = c.strikes.as<double>();
    L.slots = c.slots.as<double2>();
    L.path_payoff = dpay;
    L.path_state = dstate;
    L.fp32 = o->precision == OPE_PRECISION_FP32_PATH;
    L.code = config_code(d);
    HIP_TRY(hipEventRecord(c.ev0, s));
    HIP_TRY(launch_replay(L, s));
    HIP_TRY(hipEventRecord(c.evk, s));
    HIP_TRY(launch_reduce(c.descs.as<McDesc>(), c.jobs.as<uint2>(), 1, c.slots.as<double2>(), c.out.as<double2>(), s));
    HIP_TRY(hipEventRecord(c.ev1, s));
    double2 sums;
    HIP_TRY(hipMemcpyAsync(&sums, c.out.p, sizeof sums, hipMemcpyDeviceToHost, s));
    if (dpay) HIP_TRY(hipMemcpyAsync(o->path_payoffs, dpay, (size_t)d.n_pairs * 2 * 8, hipMemcpyDeviceToHost, s));
    if (dstate) HIP_TRY(hipMemcpyAsync(o->path_states, dstate, (size_t)d.n_pairs * 8 * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    float ms = 0.0f, kms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    HIP_TRY(hipEventElapsedTime(&kms, c.ev0, c.evk));
    if (o->device_ms) *o->device_ms = ms;
    if (o->kernel_ms) *o->kernel_ms = kms;
    finish(d, sums, out, o->stats);
    return OPE_OK;
}

int price_single(const Request& r, const ope_mc_options* o, ope_mc_result* out) {
    if (!out) return fail(OPE_ERR_INVALID_ARGUMENT, "null output");
    int rc = check_options(o);
    if (rc) return rc;
    Job job;
    if ((rc = make_desc(r, job.d))) return rc;
    job.d.flags = o ? o->flags : 0u;
    job.strikes.push_back(r.payoff.strike);
    std::lock_guard<std::mutex> lk(g_ctx.mu);
    if ((rc = ensure_init(g_ctx))) return rc;
    if (o && o->rng == OPE_RNG_REPLAY) return price_replay(g_ctx, job, o, out);
    std::vector<Job> jobs{job};
    Plan p;
    if ((rc = plan(jobs, p))) return rc;
    RawOut raw;
    if ((rc = run_plan(g_ctx, p, o && o->precision == OPE_PRECISION_FP32_PATH, raw))) return rc;
    finish(p.descs[0], raw.sums[0], out, o ? o->stats : nullptr);
    report_times(o, raw);
    return OPE_OK;
}

Request make_request(double spot, double expiry, double rate, double div_yield, const ope_heston_params* hp,
                     bool bates, double lambda, double mu_j, double sigma_j, const ope_payoff* payoff,
                     const ope_mc_config* cfg) {
    Request r;
    r.spot = spot;
    r.expiry = expiry;
    r.rate = rate;
    r.div_yield = div_yield;
    r.hp = *hp;
    r.has_jump_params = bates;
    r.lambda = lambda;
    r.mu_j = mu_j;
    r.sigma_j = sigma_j;
    r.payoff = *payoff;
    r.cfg = *cfg;
    return r;
}

int device_available() {
    std::lock_guard<std::mutex> lk(g_ctx.mu);
    return ensure_init(g_ctx);
}

int with_context(int (*body)(void* arg, const ContextView& view), void* arg) {
    std::lock_guard<std::mutex> lk(g_ctx.mu);
    const int rc = ensure_init(g_ctx);
    if (rc) return rc;
    const ContextView v{g_ctx.stream, g_ctx.ev0, g_ctx.evk, g_ctx.ev1};
    return body(arg, v);
}

}  // namespace host
}  // namespace ope

using namespace ope;
using namespace ope::host;

extern "C" {

void ope_mc_config_default(ope_mc_config* cfg) {
    if (!cfg) return;
    std::memset(cfg, 0, sizeof *cfg);
    cfg->n_paths = 200000;
    cfg->n_steps = 252;
    cfg->seed = 0xC0FFEE;
    cfg->antithetic = 1;
    cfg->scheme = OPE_SCHEME_FULL_TRUNCATION_EULER;
}

int ope_mc_heston_ex(double spot, double expiry, double rate, double div_yield, const ope_heston_params* params,
                     const ope_payoff* payoff, const ope_mc_config* cfg, const ope_mc_options* opts,
                     ope_mc_result* out) {
    if (!params || !payoff || !cfg) return fail(OPE_ERR_INVALID_ARGUMENT, "null argument");
    return price_single(make_request(spot, expiry, rate, div_yield, params, false, 0, 0, 0, payoff, cfg), opts, out);
}

int ope_mc_bates_ex(double spot, double expiry, double rate, double div_yield, const ope_heston_params* params,
                    double lambda, double mu_j, double sigma_j, const ope_payoff* payoff, const ope_mc_config* cfg,
                    const ope_mc_options* opts, ope_mc_result* out) {
    if (!params || !payoff || !cfg) return fail(OPE_ERR_INVALID_ARGUMENT, "null argument");
    return price_single(make_request(spot, expiry, rate, div_yield, params, true, lambda, mu_j, sigma_j, payoff, cfg),
                        opts, out);
}

int ope_mc_heston(double spot, double expiry, double rate, double div_yield, const ope_heston_params* params,
                  const ope_payoff* payoff, const ope_mc_config* cfg, ope_mc_result* out) {
    return ope_mc_heston_ex(spot, expiry, rate, div_yield, params, payoff, cfg, nullptr, out);
}

int ope_mc_bates(double spot, double expiry, double rate, double div_yield, const ope_heston_params* params,
                 double lambda, double mu_j, double sigma_j, const ope_payoff* payoff, const ope_mc_config* cfg,
                 ope_mc_result* out) {
    return ope_mc_bates_ex(spot, expiry, rate, div_yield, params, lambda, mu_j, sigma_j, payoff, cfg, nullptr, out);
}

int ope_mc_surface(double spot, double rate, double div_yield, const ope_heston_params* params, double lambda,
                   double mu_j, double sigma_j, int32_t opt_type, const double* expiries, uint32_t n_expiries,
                   const double* strikes, uint32_t n_strikes, const ope_mc_config* cfg, const ope_mc_options* opts,
                   ope_mc_result* out) {
    if (!params || !cfg || !expiries || !strikes || !out) return fail(OPE_ERR_INVALID_ARGUMENT, "null argument");
    if (n_expiries < 1 || n_strikes < 1) return fail(OPE_ERR_INVALID_ARGUMENT, "empty surface");
    int rc = check_options(opts);
    if (rc) return rc;
    if (opts && opts->rng == OPE_RNG_REPLAY) return fail(OPE_ERR_INVALID_ARGUMENT, "surface does not take replay draws");
    std::vector<Job> jobs(n_expiries);
    for (uint32_t e = 0; e < n_expiries; ++e) {
        ope_payoff pay{OPE_PAYOFF_EUROPEAN, opt_type, strikes[0], 0.0, 0.0};
        Request r = make_request(spot, expiries[e], rate, div_yield, params, true, lambda, mu_j, sigma_j, &pay, cfg);
        if ((rc = make_desc(r, jobs[e].d))) return rc;
        jobs[e].strikes.assign(strikes, strikes + n_strikes);
    }
    std::lock_guard<std::mutex> lk(g_ctx.mu);
    if ((rc = ensure_init(g_ctx))) return rc;
    Plan p;
    if ((rc = plan(jobs, p))) return rc;
    RawOut raw;
    if ((rc = run_plan(g_ctx, p, opts && opts->precision == OPE_PRECISION_FP32_PATH, raw))) return rc;
    for (uint32_t e = 0; e < n_expiries; ++e)
        for (uint32_t k = 0; k < n_strikes; ++k) {
            const McDesc& d = p.descs[e];
            finish(d, raw.sums[d.out_off + k], &out[(size_t)e * n_strikes + k],
                   opts && opts->stats ? &opts->stats[(size_t)e * n_strikes + k] : nullptr);
        }
    report_times(opts, raw);
    return OPE_OK;
}

int ope_mc_batch(const ope_mc_request* reqs, uint32_t n, const ope_mc_options* opts, ope_mc_result* out) {
    if (!reqs || !out) return fail(OPE_ERR_INVALID_ARGUMENT, "null argument");
    if (n == 0) return OPE_OK;
    int rc = check_options(opts);
    if (rc) return rc;
    if (opts && opts->rng == OPE_RNG_REPLAY) return fail(OPE_ERR_INVALID_ARGUMENT, "batch does not take replay draws");
    std::vector<Job> jobs(n);
    for (uint32_t i = 0; i < n; ++i) {
        const ope_mc_request& q = reqs[i];
        Request r = make_request(q.spot, q.expiry, q.rate, q.div_yield, &q.heston, true, q.lambda, q.mu_j, q.sigma_j,
                                 &q.payoff, &q.cfg);
        if ((rc = make_desc(r, jobs[i].d))) return fail(rc, "request %u: %s", i, g_last_error.c_str());
        jobs[i].d.flags = opts ? opts->flags : 0u;
        jobs[i].strikes.push_back(q.payoff.strike);
    }
    std::lock_guard<std::mutex> lk(g_ctx.mu);
    if ((rc = ensure_init(g_ctx))) return rc;
    Plan p;
    if ((rc = plan(jobs, p))) return rc;
    RawOut raw;
    if ((rc = run_plan(g_ctx, p, opts && opts->precision == OPE_PRECISION_FP32_PATH, raw))) return rc;
    for (uint32_t i = 0; i < n; ++i)
        finish(p.descs[i], raw.sums[p.descs[i].out_off], &out[i], opts && opts->stats ? &opts->stats[i] : nullptr);
    report_times(opts, raw);
    return OPE_OK;
}

int ope_get_device_info(ope_device_info* info) {
    if (!info) return fail(OPE_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(g_ctx.mu);
    int rc = ensure_init(g_ctx);
    if (rc) return rc;
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, 0));
    std::memset(info, 0, sizeof *info);
    std::snprintf(info->name, sizeof info->name, "%s", prop.name);
    std::snprintf(info->arch, sizeof info->arch, "%s", prop.gcnArchName);
    info->compute_units = prop.multiProcessorCount;
    info->clock_mhz = prop.clockRate / 1000;
    info->total_mem_bytes = prop.totalGlobalMem;
    return OPE_OK;
}

const char* ope_status_string(int status) {
    switch (status) {
    case OPE_OK: return "ok";
    case OPE_ERR_INVALID_ARGUMENT: return "invalid argument";
    case OPE_ERR_NO_DEVICE: return "no usable GPU";
    case OPE_ERR_HIP: return "HIP runtime error";
    case OPE_ERR_OUT_OF_MEMORY: return "device out of memory";
    case OPE_ERR_OVER_BUDGET: return "over device-memory budget";
    case OPE_ERR_NOT_FOUND: return "unknown ticket";
    case OPE_ERR_SHUTDOWN: return "server shut down";
    case OPE_ERR_CANCELLED: return "request cancelled";
    case OPE_ERR_DEADLINE: return "deadline passed before admission";
    case OPE_ERR_TIMEOUT: return "serving watchdog timeout";
    default: return "unknown status";
    }
}

const char* ope_last_error_message(void) { return g_last_error.c_str(); }

int ope_debug_philox(uint64_t seed, const uint32_t* ctr, uint32_t n, uint32_t* out) {
    if ((!ctr || !out) && n) return fail(OPE_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(g_ctx.mu);
    int rc = ensure_init(g_ctx);
    if (rc) return rc;
    Context& c = g_ctx;
    const size_t bytes = (size_t)n * 16;
    HIP_TRY(c.dbg_in.ensure(bytes));
    HIP_TRY(c.dbg_out.ensure(bytes));
    if (n) HIP_TRY(hipMemcpyAsync(c.dbg_in.p, ctr, bytes, hipMemcpyHostToDevice, c.stream));
    HIP_TRY(launch_debug_philox(seed, c.dbg_in.as<uint4>(), c.dbg_out.as<uint4>(), n, c.stream));
    if (n) HIP_TRY(hipMemcpyAsync(out, c.dbg_out.p, bytes, hipMemcpyDeviceToHost, c.stream));
    HIP_TRY(hipStreamSynchronize(c.stream));
    return OPE_OK;
}

int ope_debug_fp64_math(int32_t fn, const double* x, uint32_t n, double* out) {
    if ((!x || !out) && n) return fail(OPE_ERR_INVALID_ARGUMENT, "null argument");
    if (fn < 0 || fn > 7) return fail(OPE_ERR_INVALID_ARGUMENT, "unknown function %d", fn);
    if (fn == 6)
        for (uint32_t i = 0; i < n; ++i)
            if (!(x[i] >= 0.0 && x[i] <= 1.79769313486231570e308))
                return fail(OPE_ERR_INVALID_ARGUMENT, "square root argument %u is negative or not finite", i);
    if (fn == 7)
        for (uint32_t i = 0; i < n; ++i)
            if (!(std::fabs(x[i]) <= 6.8))
                return fail(OPE_ERR_INVALID_ARGUMENT, "normal tail argument %u is outside [-6.8, 6.8]", i);
    if (fn == 2 || fn == 3)
        for (uint32_t i = 0; i < n; ++i)
            if (!(x[i] >= 0.0 && x[i] < 8388608.0 && x[i] == std::floor(x[i])))
                return fail(OPE_ERR_INVALID_ARGUMENT, "angle %u is not an integer in [0, 2^23)", i);
    std::lock_guard<std::mutex> lk(g_ctx.mu);
    int rc = ensure_init(g_ctx);
    if (rc) return rc;
    Context& c = g_ctx;
    const size_t bytes = (size_t)n * 8;
    HIP_TRY(c.dbg_in.ensure(bytes));
    HIP_TRY(c.dbg_out.ensure(bytes));
    if (n) HIP_TRY(hipMemcpyAsync(c.dbg_in.p, x, bytes, hipMemcpyHostToDevice, c.stream));
    HIP_TRY(launch_debug_fp64_math(fn, c.dbg_in.as<double>(), n, c.dbg_out.as<double>(), c.stream));
    if (n) HIP_TRY(hipMemcpyAsync(out, c.dbg_out.p, bytes, hipMemcpyDeviceToHost, c.stream));
    HIP_TRY(hipStreamSynchronize(c.stream));
    return OPE_OK;
}

int ope_debug_draws(uint64_t seed, uint64_t pair0, uint64_t n_pairs, uint64_t n_steps, double expiry, double lambda,
                    double mu_j, double sigma_j, int32_t precision, double* za, double* zb, double* jump_sum,
                    uint32_t* n_jumps) {
    if (!za || !zb || !jump_sum) return fail(OPE_ERR_INVALID_ARGUMENT, "null argument");
    if (n_pairs == 0 || n_steps == 0 || n_pairs > (1ull << 24) || n_steps > kMaxSteps)
        return fail(OPE_ERR_INVALID_ARGUMENT, "bad draw grid");
    ope_heston_params hp{0.04, 1.0, 0.04, 0.5, 0.0};
    ope_payoff pay{OPE_PAYOFF_EUROPEAN, OPE_CALL, 100.0, 0.0, 0.0};
    ope_mc_config cfg;
    ope_mc_config_default(&cfg);
    cfg.n_paths = 2;
    cfg.n_steps = n_steps;
    cfg.seed = seed;
    Request r = make_request(100.0, expiry, 0.0, 0.0, &hp, true, lambda, mu_j, sigma_j, &pay, &cfg);
    McDesc d;
    int rc = make_desc(r, d);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(g_ctx.mu);
    if ((rc = ensure_init(g_ctx))) return rc;
    Context& c = g_ctx;
    const size_t n = (size_t)n_pairs * n_steps;
    std::vector<McDesc> descs{d};
    HIP_TRY(upload(c.descs, descs, c.stream));
    HIP_TRY(c.dbg_a.ensure(n * 8));
    HIP_TRY(c.dbg_b.ensure(n * 8));
    HIP_TRY(c.dbg_j.ensure(n * 8));
    HIP_TRY(c.dbg_n.ensure(n * 4));
    HIP_TRY(launch_debug_draws(c.descs.as<McDesc>(), pair0, n_pairs, precision == OPE_PRECISION_FP32_PATH,
                               c.dbg_a.as<double>(), c.dbg_b.as<double>(), c.dbg_j.as<double>(),
                               c.dbg_n.as<uint32_t>(), c.stream));
    HIP_TRY(hipMemcpyAsync(za, c.dbg_a.p, n * 8, hipMemcpyDeviceToHost, c.stream));
    HIP_TRY(hipMemcpyAsync(zb, c.dbg_b.p, n * 8, hipMemcpyDeviceToHost, c.stream));
    HIP_TRY(hipMemcpyAsync(jump_sum, c.dbg_j.p, n * 8, hipMemcpyDeviceToHost, c.stream));
    if (n_jumps) HIP_TRY(hipMemcpyAsync(n_jumps, c.dbg_n.p, n * 4, hipMemcpyDeviceToHost, c.stream));
    HIP_TRY(hipStreamSynchronize(c.stream));
    return OPE_OK;
}

}  // extern "C"
