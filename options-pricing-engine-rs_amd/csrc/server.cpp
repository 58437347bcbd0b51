// Host side of the serving loop (ope_server_*): admission against a
// device-memory budget, a slot count and the work ring, work-item
// publication for the persistent kernel (mc_server.hip), round decisions for
// requests that stop at a target standard error, deadlines, cancellation, a
// watchdog, latency bookkeeping, kernel relaunch after idle exits.
//
// Ring safety: a request reserves ring space for every tile it may ever post
// (its cap) when it is admitted and releases it when it finishes.  Unread
// ring items are always tiles of resident requests that have not completed,
// so with the reservations summing to at most the ring's capacity the host
// never overwrites an item that a worker has not read.
//
// Threads: callers submit / poll / wait under one mutex; one scheduler thread
// owns the slots, the ring and the kernel.  Host <-> device communication is
// one-directional per word (mc_server.h): the host only stores to the
// pinned control block, ring and records; the device only stores to the
// pinned completion records.
#include <hip/hip_runtime.h>
#include <sys/prctl.h>

#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <mutex>
#include <queue>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine.h"
#include "mc_launch.h"
#include "mc_server.h"

using namespace ope;
using namespace ope::host;
using Clock = std::chrono::steady_clock;

namespace {

constexpr uint32_t kRingDefault = 1u << 22;        // work items (64 MiB pinned)
constexpr uint32_t kRingMax = 1u << 26;            // 1 GiB pinned
constexpr uint32_t kTilePairs = (uint32_t)kBlock;  // one pair per thread in the server
constexpr double kSleepPollUs = 1.5;               // rough duration of one idle poll on the device
constexpr double kWatchdogDefaultMs = 30000.0;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Ticket {
    McDesc d;
    double strike = 0.0;
    double target = 0.0;
    int state = 0;  // 0 queued, 1 admitted, 2 finished, 3 failed
    int error = 0;
    bool cancel = false;        // admitted and cancelled: fails at its round's end
    bool has_deadline = false;
    uint32_t cap_tiles = 0;
    uint64_t footprint = 0;
    ope_mc_result res{};
    ope_server_request_info info{};
    Clock::time_point t_submit, t_admit, deadline;
};

using DeadlineEntry = std::pair<Clock::time_point, uint64_t>;

struct SlotState {
    bool active = false;
    uint64_t ticket = 0;
    uint64_t cum = 0;          // done[slot] once every posted tile has finished
    uint32_t posted = 0;       // tiles of the current request posted so far
    uint32_t round_tiles = 0;  // tiles of the round in flight
    uint32_t rounds = 0;
    uint64_t pool_off = 0;
};

// First-fit allocator over the partial-sum pool (units: slots of 16 bytes).
class RangeAlloc {
  public:
    void init(uint64_t n) {
        free_.clear();
        free_[0] = n;
    }
    bool alloc(uint64_t len, uint64_t* off) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= len) {
                *off = it->first;
                const uint64_t rest = it->second - len, start = it->first + len;
                free_.erase(it);
                if (rest) free_[start] = rest;
                return true;
            }
        }
        return false;
    }
    void release(uint64_t off, uint64_t len) {
        auto it = free_.emplace(off, len).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

  private:
    std::map<uint64_t, uint64_t> free_;
};

}  // namespace

struct ope_server {
    ope_server_config cfg{};
    int fp32 = 0;
    hipStream_t stream = nullptr;
    // pinned host memory
    SrvHostCtrl* hc = nullptr;
    SrvItem* ring = nullptr;
    SrvRecord* recs = nullptr;
    SrvCompletion* comps = nullptr;
    // device memory
    SrvDevCtrl* dc = nullptr;
    void* mirror = nullptr;
    uint32_t* done = nullptr;
    double2* pool = nullptr;
    uint64_t pool_slots = 0;
    SrvLaunch L{};

    uint32_t ring_cap = kRingDefault;

    std::thread th;
    std::mutex mu;
    std::condition_variable cv;      // results
    std::condition_variable wake;    // scheduler
    bool stopping = false;
    std::deque<uint64_t> queue;
    std::unordered_map<uint64_t, Ticket> tickets;
    // queued requests with a deadline, earliest first (entries of requests
    // admitted or gone meanwhile are skipped when they surface)
    std::priority_queue<DeadlineEntry, std::vector<DeadlineEntry>, std::greater<DeadlineEntry>> deadlines;
    uint64_t next_ticket = 1;
    ope_server_stats stats{};
    std::atomic<int> fatal{0};

    // scheduler-only state
    std::vector<SlotState> slots;
    std::vector<uint32_t> free_slots;
    RangeAlloc alloc;
    uint64_t tail = 0;
    uint64_t in_flight = 0;      // tiles posted and not yet reported complete
    uint64_t ring_reserved = 0;  // sum of the caps of the admitted requests
    uint64_t next_comp = 0;
    uint64_t resident_bytes = 0;
    uint32_t active = 0;
    Clock::time_point last_progress;  // last round completion, launch, or post into an idle ring

    int init();
    void run();
    bool kernel_running();
    int launch();
    void post_round(uint32_t slot, uint32_t new_total);
    void on_completion(const SrvCompletion& c);
    bool try_admit(uint64_t id, Ticket& t);
    void finish(uint32_t slot, Ticket& t, int error);
    void expire_deadlines();
    void release_gpu();
};

int ope_server::init() {
    int rc = device_available();
    if (rc) return rc;
    HIP_TRY(hipSetDevice(0));
    HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    const uint32_t n = cfg.max_requests;
    HIP_TRY(hipHostMalloc((void**)&hc, sizeof(SrvHostCtrl), hipHostMallocCoherent));
    HIP_TRY(hipHostMalloc((void**)&ring, sizeof(SrvItem) * (size_t)ring_cap, hipHostMallocCoherent));
    HIP_TRY(hipHostMalloc((void**)&recs, sizeof(SrvRecord) * (size_t)n, hipHostMallocCoherent));
    HIP_TRY(hipHostMalloc((void**)&comps, sizeof(SrvCompletion) * (size_t)n, hipHostMallocCoherent));
    std::memset(hc, 0, sizeof(SrvHostCtrl));
    std::memset(comps, 0, sizeof(SrvCompletion) * (size_t)n);
    // the caps of the resident requests sum to at most ring_cap tiles; twice
    // that leaves the first-fit pool room against fragmentation
    pool_slots = std::min<uint64_t>(cfg.memory_budget_bytes / sizeof(double2), 2ull * ring_cap);
    HIP_TRY(hipMalloc((void**)&dc, sizeof(SrvDevCtrl)));
    HIP_TRY(hipMalloc(&mirror, server_mirror_bytes()));
    HIP_TRY(hipMalloc((void**)&done, sizeof(uint32_t) * (size_t)n));
    HIP_TRY(hipMalloc((void**)&pool, sizeof(double2) * pool_slots));
    HIP_TRY(hipMemsetAsync(dc, 0, sizeof(SrvDevCtrl), stream));
    HIP_TRY(hipMemsetAsync(mirror, 0, server_mirror_bytes(), stream));  // zero: launch_server's contract
    HIP_TRY(hipMemsetAsync(done, 0, sizeof(uint32_t) * (size_t)n, stream));
    HIP_TRY(hipStreamSynchronize(stream));

    int per_cu = 0, cus = 0;
    HIP_TRY(server_occupancy(fp32, &per_cu));
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
    if (per_cu < 1) per_cu = 1;
    L.hc = hc;
    L.ring = ring;
    L.ring_cap = ring_cap;
    L.recs = recs;
    L.dc = dc;
    L.done = done;
    L.slots = pool;
    L.comps = comps;
    L.comp_cap = n;
    const double polls = cfg.idle_exit_ms * 1000.0 / kSleepPollUs;
    L.idle_limit = (uint32_t)std::min(std::max(polls, 16.0), 4.0e9);
    L.n_blocks = (uint32_t)(per_cu * cus);  // all resident; block 0 is the sentinel
    L.fp32 = fp32;
    stats.workgroups = L.n_blocks;

    slots.assign(n, SlotState());
    for (uint32_t i = n; i-- > 0;) free_slots.push_back(i);
    alloc.init(pool_slots);
    last_progress = Clock::now();
    th = std::thread([this] { run(); });
    return OPE_OK;
}

void ope_server::release_gpu() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (pool) (void)hipFree(pool);
    if (done) (void)hipFree(done);
    if (mirror) (void)hipFree(mirror);
    if (dc) (void)hipFree(dc);
    if (comps) (void)hipHostFree(comps);
    if (recs) (void)hipHostFree(recs);
    if (ring) (void)hipHostFree(ring);
    if (hc) (void)hipHostFree(hc);
    if (stream) (void)hipStreamDestroy(stream);
    pool = nullptr;
    done = nullptr;
    mirror = nullptr;
    dc = nullptr;
    comps = nullptr;
    recs = nullptr;
    ring = nullptr;
    hc = nullptr;
    stream = nullptr;
}

bool ope_server::kernel_running() { return hipStreamQuery(stream) == hipErrorNotReady; }

int ope_server::launch() {
    // launch_server resets the launch words of the mirror on the stream first;
    // the mirrored tail and the base the last launch drained at carry over
    HIP_TRY(launch_server(L, mirror, stream));
    last_progress = Clock::now();
    std::lock_guard<std::mutex> lk(mu);
    stats.kernel_launches++;
    return OPE_OK;
}

// Publishes tiles [posted, new_total) of the request in `slot` as one round.
void ope_server::post_round(uint32_t slot, uint32_t new_total) {
    SlotState& s = slots[slot];
    const uint32_t n_new = new_total - s.posted;
    s.cum += n_new;
    const uint32_t target = (uint32_t)s.cum;  // done[slot] wraps the same way
    if (in_flight == 0) last_progress = Clock::now();
    for (uint32_t t = s.posted; t < new_total; ++t) {
        SrvItem& it = ring[tail % ring_cap];
        it.slot = slot;
        it.tile = t;
        it.target = target;
        it.n_tiles = new_total;
        ++tail;
    }
    s.posted = new_total;
    s.round_tiles = n_new;
    s.rounds++;
    in_flight += n_new;
    __atomic_store_n(&hc->tail, tail, __ATOMIC_RELEASE);
}

bool ope_server::try_admit(uint64_t id, Ticket& t) {
    if (free_slots.empty()) return false;
    if (resident_bytes + t.footprint > cfg.memory_budget_bytes) return false;
    if (ring_reserved + t.cap_tiles > ring_cap) return false;
    uint64_t off = 0;
    if (!alloc.alloc(t.cap_tiles, &off)) return false;
    const uint32_t slot = free_slots.back();
    free_slots.pop_back();
    SlotState& s = slots[slot];
    s.active = true;
    s.ticket = id;
    s.posted = 0;
    s.rounds = 0;
    s.pool_off = off;
    McDesc d = t.d;
    d.slot_off = off;
    SrvRecord& r = recs[slot];
    r.d = d;
    r.strike = t.strike;
    r.pad_ = 0.0;
    resident_bytes += t.footprint;
    ring_reserved += t.cap_tiles;
    active++;
    t.state = 1;
    t.t_admit = Clock::now();
    t.info.slot = slot;
    stats.peak_active = std::max(stats.peak_active, active);
    stats.peak_resident_bytes = std::max(stats.peak_resident_bytes, resident_bytes);
    uint32_t first = t.cap_tiles;
    if (t.target > 0.0)
        first = std::min<uint32_t>(t.cap_tiles, std::max<uint32_t>(1, (cfg.initial_round_pairs + kTilePairs - 1) / kTilePairs));
    post_round(slot, first);
    return true;
}

// Releases an admitted request's slot, pool range and reservations and
// publishes its outcome (error 0: finished with t.res / t.info filled in).
// Called with mu held.
void ope_server::finish(uint32_t slot, Ticket& t, int error) {
    SlotState& s = slots[slot];
    const auto now = Clock::now();
    t.info.rounds = s.rounds;
    t.info.queue_ms = ms_between(t.t_submit, t.t_admit);
    t.info.latency_ms = ms_between(t.t_submit, now);
    if (error) {
        t.state = 3;
        t.error = error;
    } else {
        t.state = 2;
        stats.completed++;
        stats.path_steps += t.info.path_steps;
    }
    resident_bytes -= t.footprint;
    ring_reserved -= t.cap_tiles;
    active--;
    alloc.release(s.pool_off, t.cap_tiles);
    s.active = false;
    free_slots.push_back(slot);
}

void ope_server::on_completion(const SrvCompletion& c) {
    const uint32_t slot = c.slot;
    if (slot >= slots.size() || !slots[slot].active) return;
    SlotState& s = slots[slot];
    in_flight -= s.round_tiles;
    s.round_tiles = 0;
    last_progress = Clock::now();
    if (fatal.load()) return;  // every request has already failed
    std::unique_lock<std::mutex> lk(mu);
    auto found = tickets.find(s.ticket);
    if (found == tickets.end()) return;
    Ticket& t = found->second;
    if (t.cancel) {
        finish(slot, t, OPE_ERR_CANCELLED);
        lk.unlock();
        cv.notify_all();
        return;
    }
    const uint64_t per_pair = t.d.antithetic ? 2 : 1;
    const uint64_t pairs = std::min<uint64_t>((uint64_t)c.n_tiles * kTilePairs, t.d.n_pairs);
    const uint64_t count = pairs * per_pair;
    // mc.rs:198-204
    const double n = (double)count;
    const double mean = c.sum / n;
    const double var = std::max(c.sum_sq / n - mean * mean, 0.0);
    const double se = std::sqrt(var / n);
    int reason = -1;
    if (t.target > 0.0 && se <= t.target * std::fabs(mean)) reason = OPE_SERVER_STOP_TARGET;
    else if (t.target <= 0.0 || c.n_tiles >= t.cap_tiles) reason = OPE_SERVER_STOP_ALL_PATHS;
    else if (t.has_deadline && last_progress >= t.deadline) reason = OPE_SERVER_STOP_DEADLINE;
    if (reason < 0) {
        uint32_t next = t.cap_tiles;
        if (mean != 0.0) {
            const double need = (double)pairs * std::pow(se / (t.target * std::fabs(mean)), 2.0) * 1.05;
            next = (uint32_t)std::min<double>(t.cap_tiles, std::ceil(need / kTilePairs));
            next = std::max<uint32_t>(next, c.n_tiles + 1);
            next = std::min<uint32_t>(next, t.cap_tiles);
        }
        lk.unlock();
        post_round(slot, next);
        return;
    }
    t.res.price = mean;
    t.res.std_error = se;
    t.info.paths = count;
    t.info.path_steps = count * (uint64_t)t.d.n_steps;
    t.info.stop_reason = reason;
    finish(slot, t, 0);
    lk.unlock();
    cv.notify_all();
}

// Fails the queued requests whose deadline has passed.
void ope_server::expire_deadlines() {
    const auto now = Clock::now();
    bool any = false;
    std::lock_guard<std::mutex> lk(mu);
    while (!deadlines.empty() && deadlines.top().first <= now) {
        const uint64_t id = deadlines.top().second;
        deadlines.pop();
        auto it = tickets.find(id);
        if (it == tickets.end() || it->second.state != 0) continue;
        it->second.state = 3;
        it->second.error = OPE_ERR_DEADLINE;
        it->second.info.queue_ms = ms_between(it->second.t_submit, now);
        it->second.info.latency_ms = it->second.info.queue_ms;
        any = true;
    }
    if (any) cv.notify_all();
}

void ope_server::run() {
    // The loop sleeps 5 us between polls while work is outstanding; with the
    // default 50 us timer slack each sleep lasted ~55 us, which added that
    // much to a request's latency at every round boundary and admission.
    prctl(PR_SET_TIMERSLACK, 1000UL);
    auto fail_all = [this](int code) {
        {
            std::lock_guard<std::mutex> lk(mu);
            fatal.store(code);
            for (auto& kv : tickets)
                if (kv.second.state < 2) {
                    kv.second.state = 3;
                    kv.second.error = code;
                }
            queue.clear();
        }
        cv.notify_all();
    };
    for (;;) {
        bool progress = false;
        // 1. round completions
        for (;;) {
            const SrvCompletion* c = comps + (next_comp % cfg.max_requests);
            if (__atomic_load_n(&c->seq, __ATOMIC_ACQUIRE) != next_comp + 1) break;
            SrvCompletion rec;
            std::memcpy(&rec, c, sizeof rec);
            ++next_comp;
            on_completion(rec);
            progress = true;
        }
        // 2. deadlines of queued requests, then admissions, in arrival order
        expire_deadlines();
        bool stop_now = false;
        {
            std::unique_lock<std::mutex> lk(mu);
            while (!queue.empty() && fatal.load() == 0) {
                auto it = tickets.find(queue.front());
                if (it == tickets.end() || it->second.state != 0) {
                    queue.pop_front();
                    continue;
                }
                if (!try_admit(it->first, it->second)) break;
                queue.pop_front();
                progress = true;
            }
            if (stopping) {
                for (uint64_t id : queue) {
                    auto it = tickets.find(id);
                    if (it != tickets.end() && it->second.state == 0) {
                        it->second.state = 3;
                        it->second.error = OPE_ERR_SHUTDOWN;
                    }
                }
                queue.clear();
                cv.notify_all();
                stop_now = active == 0;
            }
        }
        if (stop_now) break;
        // 3. keep a kernel running while tiles are outstanding
        if (in_flight > 0 && fatal.load() == 0 && !kernel_running()) {
            const hipError_t q = hipStreamQuery(stream);
            if ((q != hipSuccess && q != hipErrorNotReady) || launch() != OPE_OK) fail_all(OPE_ERR_HIP);
            progress = true;
        }
        // 4. watchdog: tiles outstanding and no round has completed for too long.
        //    The kernel is told to drain (the published tiles still run) and
        //    every outstanding request fails.
        if (in_flight > 0 && fatal.load() == 0 && cfg.watchdog_ms > 0.0 &&
            ms_between(last_progress, Clock::now()) > cfg.watchdog_ms) {
            fail_all(OPE_ERR_TIMEOUT);
            __atomic_store_n(&hc->shutdown, 1u, __ATOMIC_RELEASE);
        }
        if (fatal.load()) {
            std::unique_lock<std::mutex> lk(mu);
            if (stopping) break;
            wake.wait_for(lk, std::chrono::milliseconds(5));
            continue;
        }
        if (!progress) {
            std::unique_lock<std::mutex> lk(mu);
            if (active == 0 && queue.empty() && !stopping)
                wake.wait_for(lk, std::chrono::milliseconds(1));
            else {
                lk.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(5));
            }
        }
    }
    __atomic_store_n(&hc->shutdown, 1u, __ATOMIC_RELEASE);
    (void)hipStreamSynchronize(stream);
}

extern "C" {

int ope_server_create(const ope_server_config* cfg, ope_server** out) {
    if (!out) return fail(OPE_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    ope_server_config c{};
    if (cfg) c = *cfg;
    if (c.precision != OPE_PRECISION_FP64 && c.precision != OPE_PRECISION_FP32_PATH)
        return fail(OPE_ERR_INVALID_ARGUMENT, "unknown precision %d", c.precision);
    if (c.max_requests == 0) c.max_requests = 1024;
    if (c.memory_budget_bytes == 0) c.memory_budget_bytes = 64ull << 30;
    if (c.initial_round_pairs == 0) c.initial_round_pairs = 65536;
    if (c.ring_items == 0) c.ring_items = kRingDefault;
    if (!(c.idle_exit_ms > 0.0)) c.idle_exit_ms = 200.0;
    if (std::isnan(c.watchdog_ms)) return fail(OPE_ERR_INVALID_ARGUMENT, "watchdog_ms is NaN");
    if (c.watchdog_ms == 0.0) c.watchdog_ms = kWatchdogDefaultMs;
    if (c.max_requests > (1u << 20)) return fail(OPE_ERR_INVALID_ARGUMENT, "max_requests above 2^20");
    if (c.ring_items > kRingMax) return fail(OPE_ERR_INVALID_ARGUMENT, "ring_items above 2^26");
    ope_server* s = new ope_server();
    s->cfg = c;
    s->ring_cap = c.ring_items;
    s->fp32 = c.precision == OPE_PRECISION_FP32_PATH;
    const int rc = s->init();
    if (rc) {
        s->release_gpu();
        delete s;
        return rc;
    }
    *out = s;
    return OPE_OK;
}

int ope_server_submit_ex(ope_server* s, const ope_mc_request* q, double target_rel_se, double deadline_ms,
                         uint64_t* ticket) {
    if (!s || !q || !ticket) return fail(OPE_ERR_INVALID_ARGUMENT, "null argument");
    if (!std::isfinite(target_rel_se)) return fail(OPE_ERR_INVALID_ARGUMENT, "non-finite target");
    if (std::isnan(deadline_ms)) return fail(OPE_ERR_INVALID_ARGUMENT, "deadline is NaN");
    Ticket t;
    Request r = make_request(q->spot, q->expiry, q->rate, q->div_yield, &q->heston, true, q->lambda, q->mu_j,
                             q->sigma_j, &q->payoff, &q->cfg);
    int rc = make_desc(r, t.d);
    if (rc) return rc;
    t.d.ppt = 1;
    t.d.n_strikes = 1;
    t.d.strike_off = 0;
    t.d.out_off = 0;
    const uint64_t cap_tiles = (t.d.n_pairs + kTilePairs - 1) / kTilePairs;
    if (cap_tiles > s->ring_cap)
        return fail(OPE_ERR_INVALID_ARGUMENT, "request of %llu tiles exceeds the serving ring (%u tiles)",
                    (unsigned long long)cap_tiles, s->ring_cap);
    t.d.n_tiles = (uint32_t)cap_tiles;
    t.cap_tiles = (uint32_t)cap_tiles;
    t.strike = q->payoff.strike;
    t.target = target_rel_se;
    t.footprint = cap_tiles * (sizeof(double2) + sizeof(SrvItem)) + sizeof(SrvRecord) + sizeof(SrvCompletion);
    if (t.footprint > s->cfg.memory_budget_bytes || cap_tiles > s->pool_slots)
        return fail(OPE_ERR_OVER_BUDGET, "request needs %llu bytes, budget %llu", (unsigned long long)t.footprint,
                    (unsigned long long)s->cfg.memory_budget_bytes);
    t.t_submit = Clock::now();
    if (deadline_ms > 0.0) {
        t.has_deadline = true;
        t.deadline = t.t_submit + std::chrono::duration_cast<Clock::duration>(
                                      std::chrono::duration<double, std::milli>(std::min(deadline_ms, 1e12)));
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stopping) return fail(OPE_ERR_SHUTDOWN, "server is shutting down");
        if (s->fatal.load()) return fail(s->fatal.load(), "server failed");
        const uint64_t id = s->next_ticket++;
        if (t.has_deadline) s->deadlines.emplace(t.deadline, id);
        s->tickets.emplace(id, t);
        s->queue.push_back(id);
        s->stats.submitted++;
        *ticket = id;
    }
    s->wake.notify_one();
    return OPE_OK;
}

int ope_server_submit(ope_server* s, const ope_mc_request* q, double target_rel_se, uint64_t* ticket) {
    return ope_server_submit_ex(s, q, target_rel_se, 0.0, ticket);
}

int ope_server_cancel(ope_server* s, uint64_t ticket) {
    if (!s) return fail(OPE_ERR_INVALID_ARGUMENT, "null argument");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->tickets.find(ticket);
        if (it == s->tickets.end() || it->second.state >= 2)
            return fail(OPE_ERR_NOT_FOUND, "request %llu is not queued or running", (unsigned long long)ticket);
        Ticket& t = it->second;
        if (t.state == 0) {  // still queued: the admission loop skips it
            t.state = 3;
            t.error = OPE_ERR_CANCELLED;
            t.info.queue_ms = t.info.latency_ms = ms_between(t.t_submit, Clock::now());
        } else {
            t.cancel = true;  // admitted: fails when its round completes
        }
    }
    s->cv.notify_all();
    return OPE_OK;
}

static int copy_out(const Ticket& t, ope_mc_result* out, ope_server_request_info* info) {
    if (t.state == 3) return fail(t.error ? t.error : OPE_ERR_HIP, "request failed");
    if (out) *out = t.res;
    if (info) *info = t.info;
    return OPE_OK;
}

int ope_server_poll(ope_server* s, uint64_t ticket, ope_mc_result* out, ope_server_request_info* info,
                    int32_t* done) {
    if (!s || !done) return fail(OPE_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return fail(OPE_ERR_NOT_FOUND, "unknown ticket %llu", (unsigned long long)ticket);
    *done = it->second.state >= 2;
    if (!*done) return OPE_OK;
    return copy_out(it->second, out, info);
}

int ope_server_wait(ope_server* s, uint64_t ticket, double timeout_ms, ope_mc_result* out,
                    ope_server_request_info* info) {
    if (!s) return fail(OPE_ERR_INVALID_ARGUMENT, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto finished = [&] {
        auto it = s->tickets.find(ticket);
        return it == s->tickets.end() || it->second.state >= 2;
    };
    if (timeout_ms < 0)
        s->cv.wait(lk, finished);
    else if (!s->cv.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), finished))
        return fail(OPE_ERR_NOT_FOUND, "request %llu not finished", (unsigned long long)ticket);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return fail(OPE_ERR_NOT_FOUND, "unknown ticket %llu", (unsigned long long)ticket);
    const int rc = copy_out(it->second, out, info);
    s->tickets.erase(it);
    return rc;
}

int ope_server_get_stats(ope_server* s, ope_server_stats* out) {
    if (!s || !out) return fail(OPE_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    return OPE_OK;
}

int ope_server_destroy(ope_server* s) {
    if (!s) return OPE_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
    }
    s->wake.notify_all();
    if (s->th.joinable()) s->th.join();
    s->cv.notify_all();
    s->release_gpu();
    delete s;
    return OPE_OK;
}

}  // extern "C"
