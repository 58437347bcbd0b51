// C ABI of the analytic pricers (include/ope_cf.h): validation, descriptors,
// one launch per call of k_cf (cf_kernels.hip) or of k_bsm / k_iv
// (bsm_kernels.hip).  The single-option entry points are batches of one.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/ope_calib.h"
#include "../../include/ope_cf.h"
#include "cf_desc.h"
#include "engine.h"

namespace ope {
hipError_t launch_cf(const CfDesc* opts, uint32_t n, double* out, CfStats* stats, double* dout, hipStream_t s);
hipError_t launch_bsm(const ope_option_contract* c, uint32_t n, double* price, ope_pricing_result* greeks,
                      hipStream_t s);
hipError_t launch_iv(const ope_option_contract* c, const double* mkt, uint32_t n, double* vol, uint8_t* found,
                     hipStream_t s);
}

using namespace ope;
using namespace ope::host;

namespace {

struct CfContext {
    std::mutex mu;
    bool ready = false;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    void* buf = nullptr;  // descriptors, then prices, then stats
    size_t cap = 0;
    void* host = nullptr;  // pinned staging laid out as buf
    size_t host_cap = 0;
};

CfContext g_cf;

constexpr size_t kMaxOptions = 1u << 24;  // one launch: one workgroup per option

int cf_init(CfContext& c) {
    if (c.ready) return OPE_OK;
    if (int rc = device_available()) return rc;
    HIP_TRY(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&c.ev0));
    HIP_TRY(hipEventCreate(&c.ev1));
    c.ready = true;
    return OPE_OK;
}

int make_cf_desc(const ope_option_contract& k, const ope_bates_params& bp, bool bates, CfDesc& d) {
    const ope_heston_params& h = bp.heston;
    for (double v : {k.spot, k.strike, k.expiry, k.rate, k.div_yield, h.v0, h.kappa, h.theta, h.sigma, h.rho})
        if (!std::isfinite(v)) return fail(OPE_ERR_INVALID_ARGUMENT, "non-finite argument");
    if (bates && !(std::isfinite(bp.lambda) && std::isfinite(bp.mu_j) && std::isfinite(bp.sigma_j)))
        return fail(OPE_ERR_INVALID_ARGUMENT, "non-finite jump parameters");
    if (!(k.spot > 0.0) || !(k.strike > 0.0))
        return fail(OPE_ERR_INVALID_ARGUMENT, "spot and strike must be > 0 (got %g, %g)", k.spot, k.strike);
    if (k.opt_type != OPE_CALL && k.opt_type != OPE_PUT)
        return fail(OPE_ERR_INVALID_ARGUMENT, "unknown option type %d", k.opt_type);
    d.spot = k.spot;
    d.strike = k.strike;
    d.expiry = k.expiry;
    d.rate = k.rate;
    d.div_yield = k.div_yield;
    d.v0 = h.v0;
    d.kappa = h.kappa;
    d.theta = h.theta;
    d.sigma = h.sigma;
    d.rho = h.rho;
    d.lambda = bates ? bp.lambda : 0.0;
    d.mu_j = bates ? bp.mu_j : 0.0;
    d.sigma_j = bates ? bp.sigma_j : 0.0;
    d.bates = bates ? 1 : 0;
    d.put = k.opt_type == OPE_PUT ? 1 : 0;
    d.deriv = 0;
    d.pad_ = 0;
    return OPE_OK;
}

// The shared device buffer, grown to at least need bytes.  Called with g_cf.mu held.
int ensure_buf(CfContext& c, size_t need) {
    if (need <= c.cap) return OPE_OK;
    if (c.buf) HIP_TRY(hipFree(c.buf));
    c.buf = nullptr;
    c.cap = 0;
    HIP_TRY(hipMalloc(&c.buf, need));
    c.cap = need;
    return OPE_OK;
}

// The pinned staging buffer, grown to at least need bytes: copies from and
// to pinned memory run as single DMA transfers instead of the runtime's
// chunked staging of pageable memory.  Called with g_cf.mu held.
int ensure_host(CfContext& c, size_t need) {
    if (need <= c.host_cap) return OPE_OK;
    if (c.host) HIP_TRY(hipHostFree(c.host));
    c.host = nullptr;
    c.host_cap = 0;
    HIP_TRY(hipHostMalloc(&c.host, need, hipHostMallocDefault));
    c.host_cap = need;
    return OPE_OK;
}

int elapsed(CfContext& c, double* kernel_ms) {
    if (kernel_ms) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, c.ev0, c.ev1));
        *kernel_ms = ms;
    }
    return OPE_OK;
}

// Prices of descs (and, with dout, 3 jump-parameter derivatives per option,
// non-zero for Bates options with the deriv flag) in one k_cf launch.
int run_cf(const std::vector<CfDesc>& descs, double* out, uint32_t* panels, double* kernel_ms,
           double* dout = nullptr) {
    const size_t n = descs.size();
    if (n == 0) return OPE_OK;
    if (n > kMaxOptions) return fail(OPE_ERR_INVALID_ARGUMENT, "more than %zu options in one call", kMaxOptions);
    std::lock_guard<std::mutex> lk(g_cf.mu);
    if (int rc = cf_init(g_cf)) return rc;
    CfContext& c = g_cf;
    const size_t b_desc = sizeof(CfDesc) * n, b_out = sizeof(double) * n, b_stats = sizeof(CfStats) * n;
    const size_t b_dout = dout ? 3 * sizeof(double) * n : 0;
    const size_t total = b_desc + b_out + b_stats + b_dout;
    if (int rc = ensure_buf(c, total)) return rc;
    if (int rc = ensure_host(c, total)) return rc;
    char* base = static_cast<char*>(c.buf);
    char* hbase = static_cast<char*>(c.host);
    CfDesc* d_desc = reinterpret_cast<CfDesc*>(base);
    double* d_out = reinterpret_cast<double*>(base + b_desc);
    CfStats* d_stats = reinterpret_cast<CfStats*>(base + b_desc + b_out);
    double* d_dout = dout ? reinterpret_cast<double*>(base + b_desc + b_out + b_stats) : nullptr;
    std::memcpy(hbase, descs.data(), b_desc);
    HIP_TRY(hipMemcpyAsync(d_desc, hbase, b_desc, hipMemcpyHostToDevice, c.stream));
    HIP_TRY(hipEventRecord(c.ev0, c.stream));
    HIP_TRY(launch_cf(d_desc, (uint32_t)n, d_out, panels ? d_stats : nullptr, d_dout, c.stream));
    HIP_TRY(hipEventRecord(c.ev1, c.stream));
    // Prices, stats and derivatives in one transfer (the stats bytes are
    // unwritten when no panel counts were asked for, and unread then).
    HIP_TRY(hipMemcpyAsync(hbase + b_desc, d_out, total - b_desc, hipMemcpyDeviceToHost, c.stream));
    HIP_TRY(hipStreamSynchronize(c.stream));
    std::memcpy(out, hbase + b_desc, b_out);
    if (dout) std::memcpy(dout, hbase + b_desc + b_out + b_stats, b_dout);
    if (panels) {
        const CfStats* st = reinterpret_cast<const CfStats*>(hbase + b_desc + b_out);
        for (size_t i = 0; i < n; ++i) {
            panels[2 * i] = st[i].panels_p1;
            panels[2 * i + 1] = st[i].panels_p2;
        }
    }
    return elapsed(c, kernel_ms);
}

int check_contracts(const ope_option_contract* k, size_t n) {
    for (size_t i = 0; i < n; ++i) {
        for (double v : {k[i].spot, k[i].strike, k[i].expiry, k[i].rate, k[i].div_yield, k[i].vol})
            if (!std::isfinite(v)) return fail(OPE_ERR_INVALID_ARGUMENT, "option %zu: non-finite argument", i);
        if (k[i].opt_type != OPE_CALL && k[i].opt_type != OPE_PUT)
            return fail(OPE_ERR_INVALID_ARGUMENT, "option %zu: unknown option type %d", i, k[i].opt_type);
    }
    return OPE_OK;
}

constexpr size_t kMaxLanes = 1ull << 31;  // one lane per option

int run_bsm(const ope_option_contract* k, size_t n, double* price, ope_pricing_result* greeks, double* kernel_ms) {
    if (kernel_ms) *kernel_ms = 0.0;
    if (n == 0) return OPE_OK;
    if (!k || (!price && !greeks)) return fail(OPE_ERR_INVALID_ARGUMENT, "null argument");
    if (n >= kMaxLanes) return fail(OPE_ERR_INVALID_ARGUMENT, "too many options in one call");
    if (int rc = check_contracts(k, n)) return rc;
    std::lock_guard<std::mutex> lk(g_cf.mu);
    if (int rc = cf_init(g_cf)) return rc;
    CfContext& c = g_cf;
    const size_t b_in = sizeof(ope_option_contract) * n;
    const size_t b_out = greeks ? sizeof(ope_pricing_result) * n : sizeof(double) * n;
    if (int rc = ensure_buf(c, b_in + b_out)) return rc;
    char* base = static_cast<char*>(c.buf);
    auto* d_in = reinterpret_cast<ope_option_contract*>(base);
    void* d_out = base + b_in;
    HIP_TRY(hipMemcpyAsync(d_in, k, b_in, hipMemcpyHostToDevice, c.stream));
    HIP_TRY(hipEventRecord(c.ev0, c.stream));
    HIP_TRY(launch_bsm(d_in, (uint32_t)n, greeks ? nullptr : static_cast<double*>(d_out),
                       greeks ? static_cast<ope_pricing_result*>(d_out) : nullptr, c.stream));
    HIP_TRY(hipEventRecord(c.ev1, c.stream));
    HIP_TRY(hipMemcpyAsync(greeks ? static_cast<void*>(greeks) : static_cast<void*>(price), d_out, b_out,
                           hipMemcpyDeviceToHost, c.stream));
    HIP_TRY(hipStreamSynchronize(c.stream));
    return elapsed(c, kernel_ms);
}

int run_iv(const ope_option_contract* k, const double* mkt, size_t n, double* vol, uint8_t* found, double* kernel_ms) {
    if (kernel_ms) *kernel_ms = 0.0;
    if (n == 0) return OPE_OK;
    if (!k || !mkt || !vol || !found) return fail(OPE_ERR_INVALID_ARGUMENT, "null argument");
    if (n >= kMaxLanes) return fail(OPE_ERR_INVALID_ARGUMENT, "too many options in one call");
    if (int rc = check_contracts(k, n)) return rc;
    std::lock_guard<std::mutex> lk(g_cf.mu);
    if (int rc = cf_init(g_cf)) return rc;
    CfContext& c = g_cf;
    const size_t b_in = sizeof(ope_option_contract) * n, b_mkt = sizeof(double) * n, b_vol = sizeof(double) * n;
    if (int rc = ensure_buf(c, b_in + b_mkt + b_vol + n)) return rc;
    char* base = static_cast<char*>(c.buf);
    auto* d_in = reinterpret_cast<ope_option_contract*>(base);
    auto* d_mkt = reinterpret_cast<double*>(base + b_in);
    auto* d_vol = reinterpret_cast<double*>(base + b_in + b_mkt);
    auto* d_found = reinterpret_cast<uint8_t*>(base + b_in + b_mkt + b_vol);
    HIP_TRY(hipMemcpyAsync(d_in, k, b_in, hipMemcpyHostToDevice, c.stream));
    HIP_TRY(hipMemcpyAsync(d_mkt, mkt, b_mkt, hipMemcpyHostToDevice, c.stream));
    HIP_TRY(hipEventRecord(c.ev0, c.stream));
    HIP_TRY(launch_iv(d_in, d_mkt, (uint32_t)n, d_vol, d_found, c.stream));
    HIP_TRY(hipEventRecord(c.ev1, c.stream));
    HIP_TRY(hipMemcpyAsync(vol, d_vol, b_vol, hipMemcpyDeviceToHost, c.stream));
    HIP_TRY(hipMemcpyAsync(found, d_found, n, hipMemcpyDeviceToHost, c.stream));
    HIP_TRY(hipStreamSynchronize(c.stream));
    return elapsed(c, kernel_ms);
}

// greeks.rs:19-82: the 14 bumped prices of one option, in this order.
constexpr int kBumps = 14;

void bump_descs(const CfDesc& base, CfDesc* d) {
    const double ds = 0.01 * base.spot, dv = 0.01, dr = 1e-4, dt = 1.0 / 365.0;
    const double v_cur = std::sqrt(base.v0);
    const double vu = v_cur + dv, vd = std::fmax(v_cur - dv, 1e-8);
    const double vu2 = v_cur + 2.0 * dv, vd2 = std::fmax(v_cur - 2.0 * dv, 1e-8);
    for (int i = 0; i < kBumps; ++i) d[i] = base;
    d[1].spot = base.spot + ds;  // delta, gamma
    d[2].spot = base.spot - ds;
    d[3].v0 = vu * vu;  // vega
    d[4].v0 = vd * vd;
    d[5].expiry = std::fmax(base.expiry - dt, 1e-6);  // theta
    d[6].rate = base.rate + dr;  // rho
    d[7].rate = base.rate - dr;
    d[8].v0 = vu * vu;  // vanna: delta at v+dv and at v-dv
    d[8].spot = base.spot + ds;
    d[9].v0 = vu * vu;
    d[9].spot = base.spot - ds;
    d[10].v0 = vd * vd;
    d[10].spot = base.spot + ds;
    d[11].v0 = vd * vd;
    d[11].spot = base.spot - ds;
    d[12].v0 = vu2 * vu2;  // volga
    d[13].v0 = vd2 * vd2;
}

ope_pricing_result combine(const CfDesc& base, const double* p) {
    const double ds = 0.01 * base.spot, dv = 0.01, dr = 1e-4, dt = 1.0 / 365.0;
    ope_pricing_result g;
    const double price = p[0];
    g.price = price;
    g.delta = (p[1] - p[2]) / (2.0 * ds);
    g.gamma = (p[1] - 2.0 * price + p[2]) / (ds * ds);
    g.vega = (p[3] - p[4]) / (2.0 * dv);
    g.theta = (p[5] - price) / dt;
    g.rho = (p[6] - p[7]) / (2.0 * dr);
    const double delta_vup = (p[8] - p[9]) / (2.0 * ds);
    const double delta_vdn = (p[10] - p[11]) / (2.0 * ds);
    g.vanna = (delta_vup - delta_vdn) / (2.0 * dv);
    const double vega_up = (p[12] - price) / (2.0 * dv);
    const double vega_dn = (price - p[13]) / (2.0 * dv);
    g.volga = (vega_up - vega_dn) / (2.0 * dv);
    return g;
}

int run_greeks(const std::vector<CfDesc>& base, ope_pricing_result* out) {
    const size_t n = base.size();
    if (n > kMaxOptions / kBumps) return fail(OPE_ERR_INVALID_ARGUMENT, "too many options in one call");
    std::vector<CfDesc> all(n * kBumps);
    for (size_t i = 0; i < n; ++i) bump_descs(base[i], &all[i * kBumps]);
    std::vector<double> px(all.size());
    if (int rc = run_cf(all, px.data(), nullptr, nullptr)) return rc;
    for (size_t i = 0; i < n; ++i) out[i] = combine(base[i], &px[i * kBumps]);
    return OPE_OK;
}

// calibration.rs:123-126 and 162-171: bump sizes and bounds per parameter,
// in CalibModel::to_vec order.
constexpr double kCalBump[8] = {1e-4, 1e-3, 1e-4, 1e-3, 1e-4, 1e-4, 1e-4, 1e-4};
constexpr double kCalLo[8] = {1e-8, 1e-6, 1e-8, 1e-6, -0.9999, 0.0, -2.0, 1e-6};
constexpr double kCalHi[8] = {5.0, 50.0, 5.0, 10.0, 0.9999, 10.0, 2.0, 2.0};

// Residuals (and, with jac, the Jacobian) of one calibration evaluation
// (calibration.rs:573-587, 613-652).  Every price in one k_cf launch, every
// implied volatility in one k_iv launch.  Columns are central differences
// with the bump points clamped to the bounds, except Bates' three jump
// columns, which the reference takes from forward-mode AD (calibration.rs:
// 186-199, 617-634): d(price)/d(param) integrated beside the base price
// (k_cf's derivative pass) divided by the BSM vega at the model's implied
// vol; a column falls back to differences when any quote's solver finds no
// volatility or its vega is below 1e-10, as the reference's does.
int calib_eval(const ope_calib_quote* quotes, size_t n, int32_t model, const double* params, double* res,
               double* jac) {
    if (model != OPE_CF_HESTON && model != OPE_CF_BATES)
        return fail(OPE_ERR_INVALID_ARGUMENT, "unknown model %d", model);
    const bool bates = model == OPE_CF_BATES;
    const int dim = bates ? 8 : 5;
    for (int c = 0; c < dim; ++c)
        if (!std::isfinite(params[c])) return fail(OPE_ERR_INVALID_ARGUMENT, "non-finite parameter %d", c);
    std::vector<double> iv_mkt(n);
    for (size_t i = 0; i < n; ++i) {
        iv_mkt[i] = quotes[i].iv_market;
        if (!std::isfinite(quotes[i].iv_market) || !std::isfinite(quotes[i].weight))
            return fail(OPE_ERR_INVALID_ARGUMENT, "quote %zu: non-finite iv_market or weight", i);
    }
    std::array<double, 8> base{};
    for (int c = 0; c < dim; ++c) base[c] = params[c];
    const bool ad = jac && bates;          // AD jump columns need the base point
    const bool with_base = res || ad;
    const int fd_cols = ad ? 5 : dim;      // columns differenced up front

    // Prices and implied vols of the base point (optional) and of the (up,
    // down) bump vectors of the given columns, in one k_cf and one k_iv
    // launch.  r(v, i) is the residual of quote i under vector v.
    struct Eval {
        std::vector<int> cols;
        std::vector<double> dist, px, iv, dpx;
        std::vector<uint8_t> found;
        size_t v0 = 0;
    };
    auto evaluate = [&](bool want_base, const std::vector<int>& cols, Eval& e) -> int {
        std::vector<std::array<double, 8>> vecs;
        if (want_base) vecs.push_back(base);
        e.v0 = want_base ? 1 : 0;
        for (int c : cols) {
            std::array<double, 8> vu = base, vd = base;
            vu[c] = std::fmin(base[c] + kCalBump[c], kCalHi[c]);
            vd[c] = std::fmax(base[c] - kCalBump[c], kCalLo[c]);
            const double h = vu[c] - vd[c];
            if (h <= 0.0) continue;  // pinned against both bounds: zero column
            vecs.push_back(vu);
            vecs.push_back(vd);
            e.dist.push_back(h);
            e.cols.push_back(c);
        }
        const size_t nv = vecs.size();
        if (nv == 0) return OPE_OK;
        if (n > kMaxOptions / nv) return fail(OPE_ERR_INVALID_ARGUMENT, "too many quotes");
        std::vector<CfDesc> d(nv * n);
        std::vector<ope_option_contract> k(nv * n);
        for (size_t v = 0; v < nv; ++v) {
            ope_bates_params bp{};
            bp.heston = ope_heston_params{vecs[v][0], vecs[v][1], vecs[v][2], vecs[v][3], vecs[v][4]};
            bp.lambda = vecs[v][5];
            bp.mu_j = vecs[v][6];
            bp.sigma_j = vecs[v][7];
            for (size_t i = 0; i < n; ++i) {
                if (int rc = make_cf_desc(quotes[i].contract, bp, bates, d[v * n + i]))
                    return fail(rc, "quote %zu: %s", i, g_last_error.c_str());
                d[v * n + i].deriv = (ad && want_base && v == 0) ? 1 : 0;
                k[v * n + i] = quotes[i].contract;
            }
        }
        e.px.resize(nv * n);
        e.iv.resize(nv * n);
        e.found.resize(nv * n);
        if (ad && want_base) e.dpx.resize(3 * nv * n);
        if (int rc = run_cf(d, e.px.data(), nullptr, nullptr, e.dpx.empty() ? nullptr : e.dpx.data())) return rc;
        return run_iv(k.data(), e.px.data(), nv * n, e.iv.data(), e.found.data(), nullptr);
    };
    // calibration.rs:573-581: iv - iv_market, or 0 where no volatility is found
    auto r = [&](const Eval& e, size_t v, size_t i) {
        return e.found[v * n + i] ? e.iv[v * n + i] - iv_mkt[i] : 0.0;
    };
    auto fd_fill = [&](const Eval& e) {
        for (size_t p = 0; p < e.cols.size(); ++p) {
            const size_t vu = e.v0 + 2 * p, vd = vu + 1;
            for (size_t i = 0; i < n; ++i) jac[i * dim + e.cols[p]] = (r(e, vu, i) - r(e, vd, i)) / e.dist[p];
        }
    };

    std::vector<int> cols;
    if (jac)
        for (int c = 0; c < fd_cols; ++c) cols.push_back(c);
    Eval e;
    if (int rc = evaluate(with_base, cols, e)) return rc;
    if (res)
        for (size_t i = 0; i < n; ++i) res[i] = r(e, 0, i);
    if (!jac) return OPE_OK;
    for (size_t i = 0; i < n * (size_t)dim; ++i) jac[i] = 0.0;
    fd_fill(e);
    if (!ad) return OPE_OK;

    // calibration.rs:601-609 model_iv_and_vega: BSM vega at the implied vol of
    // the model price (bsm.rs:33, per unit of volatility)
    std::vector<ope_option_contract> kv(n);
    bool all_found = true;
    for (size_t i = 0; i < n; ++i) {
        kv[i] = quotes[i].contract;
        kv[i].vol = e.found[i] ? e.iv[i] : 0.2;
        all_found = all_found && e.found[i];
    }
    std::vector<ope_pricing_result> g(n);
    if (all_found)
        if (int rc = run_bsm(kv.data(), n, nullptr, g.data(), nullptr)) return rc;
    std::vector<int> fallback;
    for (int c = 5; c < 8; ++c) {
        bool ok = all_found;
        for (size_t i = 0; ok && i < n; ++i) ok = std::fabs(g[i].vega) >= 1e-10;
        if (!ok) {
            fallback.push_back(c);
            continue;
        }
        for (size_t i = 0; i < n; ++i) jac[i * dim + c] = e.dpx[3 * i + (c - 5)] / g[i].vega;
    }
    if (!fallback.empty()) {
        Eval f;
        if (int rc = evaluate(false, fallback, f)) return rc;
        fd_fill(f);
    }
    return OPE_OK;
}

ope_option_contract contract(double spot, double strike, double expiry, double rate, double div_yield,
                             int32_t opt_type) {
    ope_option_contract k{};
    k.spot = spot;
    k.strike = strike;
    k.expiry = expiry;
    k.rate = rate;
    k.div_yield = div_yield;
    k.opt_type = opt_type;
    return k;
}

}  // namespace

extern "C" {

int ope_heston_price(double spot, double strike, double expiry, double rate, double div_yield,
                     const ope_heston_params* params, int32_t opt_type, double* out) {
    if (!params || !out) return fail(OPE_ERR_INVALID_ARGUMENT, "null argument");
    ope_bates_params bp{};
    bp.heston = *params;
    std::vector<CfDesc> d(1);
    if (int rc = make_cf_desc(contract(spot, strike, expiry, rate, div_yield, opt_type), bp, false, d[0])) return rc;
    return run_cf(d, out, nullptr, nullptr);
}

int ope_bates_price(double spot, double strike, double expiry, double rate, double div_yield,
                    const ope_bates_params* params, int32_t opt_type, double* out) {
    if (!params || !out) return fail(OPE_ERR_INVALID_ARGUMENT, "null argument");
    std::vector<CfDesc> d(1);
    if (int rc = make_cf_desc(contract(spot, strike, expiry, rate, div_yield, opt_type), *params, true, d[0]))
        return rc;
    return run_cf(d, out, nullptr, nullptr);
}

int ope_batch_heston(const ope_option_contract* contracts, size_t n, const ope_heston_params* params, double* out) {
    if (n == 0) return OPE_OK;
    if (!contracts || !params || !out) return fail(OPE_ERR_INVALID_ARGUMENT, "null argument");
    ope_bates_params bp{};
    bp.heston = *params;
    std::vector<CfDesc> d(n);
    for (size_t i = 0; i < n; ++i)
        if (int rc = make_cf_desc(contracts[i], bp, false, d[i])) return fail(rc, "option %zu: %s", i, g_last_error.c_str());
    return run_cf(d, out, nullptr, nullptr);
}

int ope_batch_bates(const ope_option_contract* contracts, size_t n, const ope_bates_params* params, double* out) {
    if (n == 0) return OPE_OK;
    if (!contracts || !params || !out) return fail(OPE_ERR_INVALID_ARGUMENT, "null argument");
    std::vector<CfDesc> d(n);
    for (size_t i = 0; i < n; ++i)
        if (int rc = make_cf_desc(contracts[i], *params, true, d[i])) return fail(rc, "option %zu: %s", i, g_last_error.c_str());
    return run_cf(d, out, nullptr, nullptr);
}

int ope_heston_price_and_greeks(double spot, double strike, double expiry, double rate, double div_yield,
                                const ope_heston_params* params, int32_t opt_type, ope_pricing_result* out) {
    if (!params || !out) return fail(OPE_ERR_INVALID_ARGUMENT, "null argument");
    ope_bates_params bp{};
    bp.heston = *params;
    std::vector<CfDesc> d(1);
    if (int rc = make_cf_desc(contract(spot, strike, expiry, rate, div_yield, opt_type), bp, false, d[0])) return rc;
    return run_greeks(d, out);
}

int ope_bates_price_and_greeks(double spot, double strike, double expiry, double rate, double div_yield,
                               const ope_bates_params* params, int32_t opt_type, ope_pricing_result* out) {
    if (!params || !out) return fail(OPE_ERR_INVALID_ARGUMENT, "null argument");
    std::vector<CfDesc> d(1);
    if (int rc = make_cf_desc(contract(spot, strike, expiry, rate, div_yield, opt_type), *params, true, d[0]))
        return rc;
    return run_greeks(d, out);
}

int ope_batch_heston_greeks(const ope_option_contract* contracts, size_t n, const ope_heston_params* params,
                            ope_pricing_result* out) {
    if (n == 0) return OPE_OK;
    if (!contracts || !params || !out) return fail(OPE_ERR_INVALID_ARGUMENT, "null argument");
    ope_bates_params bp{};
    bp.heston = *params;
    std::vector<CfDesc> d(n);
    for (size_t i = 0; i < n; ++i)
        if (int rc = make_cf_desc(contracts[i], bp, false, d[i])) return fail(rc, "option %zu: %s", i, g_last_error.c_str());
    return run_greeks(d, out);
}

int ope_batch_bates_greeks(const ope_option_contract* contracts, size_t n, const ope_bates_params* params,
                           ope_pricing_result* out) {
    if (n == 0) return OPE_OK;
    if (!contracts || !params || !out) return fail(OPE_ERR_INVALID_ARGUMENT, "null argument");
    std::vector<CfDesc> d(n);
    for (size_t i = 0; i < n; ++i)
        if (int rc = make_cf_desc(contracts[i], *params, true, d[i])) return fail(rc, "option %zu: %s", i, g_last_error.c_str());
    return run_greeks(d, out);
}

int ope_bsm_price_and_greeks(const ope_option_contract* contract, ope_pricing_result* out) {
    if (!contract || !out) return fail(OPE_ERR_INVALID_ARGUMENT, "null argument");
    return run_bsm(contract, 1, nullptr, out, nullptr);
}

int ope_bsm_price(const ope_option_contract* contract, double* out) {
    if (!contract || !out) return fail(OPE_ERR_INVALID_ARGUMENT, "null argument");
    return run_bsm(contract, 1, out, nullptr, nullptr);
}

int ope_implied_vol(const ope_option_contract* contract, double market_price, double* vol, uint8_t* found) {
    return run_iv(contract, &market_price, 1, vol, found, nullptr);
}

int ope_batch_bsm(const ope_option_contract* contracts, size_t n, ope_pricing_result* out) {
    return run_bsm(contracts, n, nullptr, out, nullptr);
}

int ope_batch_bsm_price(const ope_option_contract* contracts, size_t n, double* out) {
    return run_bsm(contracts, n, out, nullptr, nullptr);
}

int ope_batch_implied_vol(const ope_option_contract* contracts, const double* market_prices, size_t n, double* vol,
                          uint8_t* found) {
    return run_iv(contracts, market_prices, n, vol, found, nullptr);
}

int ope_batch_bsm_ex(const ope_option_contract* contracts, size_t n, double* out_price, ope_pricing_result* out_greeks,
                     double* kernel_ms) {
    return run_bsm(contracts, n, out_greeks ? nullptr : out_price, out_greeks, kernel_ms);
}

int ope_batch_implied_vol_ex(const ope_option_contract* contracts, const double* market_prices, size_t n,
                             double* vol, uint8_t* found, double* kernel_ms) {
    return run_iv(contracts, market_prices, n, vol, found, kernel_ms);
}

int ope_calib_residuals(const ope_calib_quote* quotes, size_t n, int32_t model, const double* params,
                        double* residuals) {
    if (n == 0) return OPE_OK;
    if (!quotes || !params || !residuals) return fail(OPE_ERR_INVALID_ARGUMENT, "null argument");
    return calib_eval(quotes, n, model, params, residuals, nullptr);
}

int ope_calib_jacobian(const ope_calib_quote* quotes, size_t n, int32_t model, const double* params,
                       double* residuals, double* jacobian, double* jtj, double* grad) {
    if (n == 0) return OPE_OK;
    if (!quotes || !params || !jacobian) return fail(OPE_ERR_INVALID_ARGUMENT, "null argument");
    if (grad && !residuals) return fail(OPE_ERR_INVALID_ARGUMENT, "grad needs residuals");
    if (int rc = calib_eval(quotes, n, model, params, residuals, jacobian)) return rc;
    const int dim = model == OPE_CF_BATES ? 8 : 5;
    // calibration.rs:655-667
    if (jtj)
        for (int a = 0; a < dim * dim; ++a) jtj[a] = 0.0;
    if (grad)
        for (int a = 0; a < dim; ++a) grad[a] = 0.0;
    for (size_t row = 0; row < n; ++row) {
        const double w2 = quotes[row].weight * quotes[row].weight;
        const double* j = jacobian + row * dim;
        for (int c1 = 0; c1 < dim; ++c1) {
            if (grad) grad[c1] += w2 * j[c1] * residuals[row];
            if (jtj)
                for (int c2 = 0; c2 < dim; ++c2) jtj[c1 * dim + c2] += w2 * j[c1] * j[c2];
        }
    }
    return OPE_OK;
}

int ope_cf_price_options(const ope_cf_option* options, size_t n, double* out, uint32_t* panels, double* kernel_ms) {
    if (kernel_ms) *kernel_ms = 0.0;
    if (n == 0) return OPE_OK;
    if (!options || !out) return fail(OPE_ERR_INVALID_ARGUMENT, "null argument");
    std::vector<CfDesc> d(n);
    for (size_t i = 0; i < n; ++i) {
        const ope_cf_option& o = options[i];
        if (o.model != OPE_CF_HESTON && o.model != OPE_CF_BATES)
            return fail(OPE_ERR_INVALID_ARGUMENT, "option %zu: unknown model %d", i, o.model);
        if (int rc = make_cf_desc(o.contract, o.params, o.model == OPE_CF_BATES, d[i]))
            return fail(rc, "option %zu: %s", i, g_last_error.c_str());
    }
    return run_cf(d, out, panels, kernel_ms);
}

}  // extern "C"
