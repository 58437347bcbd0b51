// Persistent serving kernel: workgroups loop over work items (one tile of one
// resident request each) published by the host, until the host shuts the
// server down or the queue stays empty for idle_limit polls.
//
// * Workgroup 0 is the sentinel (dispatched first, so always resident): it
//   polls the host's tail and shutdown words (system scope) and mirrors the
//   tail into device memory, so that the other workgroups poll L2 instead of
//   the host bus.  It alone decides that the kernel drains: on host shutdown,
//   or when neither the tail nor the ticket counter has moved for idle_limit
//   polls and every item below the mirrored tail has been claimed (so an
//   idle exit never leaves an item below the tail unclaimed, even when every
//   resident worker is busy on a long tile).  Before it sets the drain flag
//   it records the last tail it mirrored as the next launch's base.
// * Each launch numbers its tickets from launch_base, which k_server_prepare
//   copies from the previous launch's next_base before the launch (stream
//   order); the sentinel never writes launch_base, so a workgroup dispatched
//   late, after the sentinel drained, still numbers its ticket correctly.
// * A free worker takes the next ticket with one fetch-and-add (tickets are
//   numbered from that base) and waits until the mirrored tail passes it: the
//   ticket is the ring index of its next work item.  Idle workers hold the
//   tickets just past the tail, so a new item starts at once on an idle
//   workgroup, and the claim costs one uncontended atomic per item however
//   many workgroups wait.  A worker leaves once draining is set and its
//   ticket is at or beyond the final tail: every item below that tail was
//   taken by a worker that runs it, and the rest wait for the next launch.
// * A worker reads the item and, when the item belongs to another request
//   than its previous one, the request's record from host memory with
//   system-scope loads (never through the scalar cache), copies the record
//   into SGPRs, and runs the same tile body as the batch kernels.
// * Completing a tile bumps done[slot]; the workgroup that brings it to the
//   item's round target reduces the request's tiles in the canonical order
//   (block_sum_slots, as k_reduce) and stores a completion record to host
//   memory, its sequence word last, with release semantics.
// Built with -ffp-contract=on, as the batch kernels.
#include "mc_device.h"
#include "mc_dispatch.h"
#include "mc_server.h"

namespace ope {
namespace {

constexpr int kRecWords = (int)(sizeof(SrvRecord) / 4);
constexpr int kRecQwords = kRecWords / 2;
static_assert(sizeof(SrvRecord) % 8 == 0, "record read in 8-byte words");

// Device-memory words shared by the sentinel and the workers.
// k_server_prepare sets launch_base = next_base and clears tickets and
// shutdown before every launch; tail and next_base carry over.
struct SrvMirror {
    uint64_t tail;         // host tail as last seen by the sentinel
    uint64_t next_base;    // tail at which the last launch drained (written by its sentinel)
    uint64_t launch_base;  // this launch's first ticket (read-only during the launch)
    uint64_t tickets;      // tickets taken in this launch
    uint32_t shutdown;     // draining
    uint32_t pad_;
};

// Polling loads are relaxed, and the acquire is one fence taken after the
// awaited value has arrived.  An acquire load invalidates the L2 each time it
// executes (buffer_inv); with a thousand idle workgroups polling, those
// invalidations queued behind each other and every memory access of the
// running tiles waited ~30 us (a one-step request took 0.14 ms end to end,
// 0.02 ms with relaxed polls).
__device__ __forceinline__ uint64_t sys_load64(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ uint32_t sys_load32(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

__device__ void sentinel(const SrvLaunch& L, SrvMirror* mirror) {
    if (threadIdx.x != 0) return;
    const uint64_t base = mirror->launch_base;
    uint64_t last = ~0ull, last_tk = 0, mirrored = base;
    uint32_t idle = 0;
    for (;;) {
        const uint64_t t = sys_load64(&L.hc->tail);
        const uint32_t stop = sys_load32(&L.hc->shutdown);
        const uint64_t tk = __hip_atomic_load(&mirror->tickets, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (t != last) {
            // pairs with the host's release store of the tail: the items below
            // it are visible to the workers that acquire the mirrored tail
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
            __hip_atomic_store(&mirror->tail, t, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            last = t;
            mirrored = t;
            idle = 0;
        } else if (tk != last_tk) {
            idle = 0;
        } else if (idle <= L.idle_limit) {
            ++idle;
        } else if (base + tk >= mirrored) {
            break;  // idle, and every item below the mirrored tail has a worker
        }
        last_tk = tk;
        // the host stops only once every published item has completed
        if (stop) break;
        __builtin_amdgcn_s_sleep(16);
    }
    __hip_atomic_store(&mirror->next_base, mirrored, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&mirror->shutdown, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

template <class Real>
__global__ __launch_bounds__(kBlock, (std::is_same<Real, float>::value ? kPinnedWavesServerF32 : 1)) void k_server(SrvLaunch L, SrvMirror* mirror) {
    __shared__ double2 lds[kTileLds];
    __shared__ uint64_t s_rec[kRecQwords];
    __shared__ SrvItem s_item;
    __shared__ uint32_t s_status;  // 0 work, 1 leave
    __shared__ uint32_t s_reload;  // the item belongs to another request than the record in s_rec
    __shared__ uint64_t s_key;     // request of the record in s_rec
    __shared__ uint32_t s_last;
    __shared__ double s_strike;
    __shared__ double2 s_w[kBlock / 64];

    if (blockIdx.x == 0) {
        sentinel(L, mirror);
        return;
    }
    // set by k_server_prepare, which ran before this launch
    const uint64_t base = __hip_atomic_load(&mirror->launch_base, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (threadIdx.x == 0) s_key = ~0ull;
    for (;;) {
        if (threadIdx.x == 0) {
            const uint64_t ticket =
                base + __hip_atomic_fetch_add(&mirror->tickets, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            uint32_t status = 0;
            SrvItem it{0, 0, 0, 0};
            for (;;) {
                const uint64_t t = __hip_atomic_load(&mirror->tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (ticket < t) {
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                    const uint64_t* src = reinterpret_cast<const uint64_t*>(L.ring + (ticket % L.ring_cap));
                    const uint64_t w0 = sys_load64(src), w1 = sys_load64(src + 1);
                    it.slot = (uint32_t)w0;
                    it.tile = (uint32_t)(w0 >> 32);
                    it.target = (uint32_t)w1;
                    it.n_tiles = (uint32_t)(w1 >> 32);
                    break;
                }
                if (__hip_atomic_load(&mirror->shutdown, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                    // the sentinel stored its final tail before the flag
                    if (ticket >= __hip_atomic_load(&mirror->tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                        status = 1;
                        break;
                    }
                    continue;
                }
                __builtin_amdgcn_s_sleep(32);
            }
            s_item = it;
            s_status = status;
            // target - n_tiles is the slot's cumulative tile count before the
            // request was admitted: constant over the request's items and
            // strictly larger for the next request in the slot, so (slot,
            // target - n_tiles) names the request, whose record does not
            // change while it runs
            const uint64_t key = ((uint64_t)it.slot << 32) | (uint32_t)(it.target - it.n_tiles);
            s_reload = key != s_key;
            s_key = key;
        }
        __syncthreads();
        if (s_status) return;
        const SrvItem it = s_item;

        // the request's record: host memory -> LDS -> SGPRs.  It is read from
        // the host only when the item belongs to another request than the
        // workgroup's previous item: host reads of one cache line from many
        // workgroups serialise (a burst of 64 requests of 3906 tiles each ran
        // at half the batch kernel's rate while every tile re-read its record
        // in 4-byte words, at 0.95 of it with the record kept and read in
        // 8-byte words)
        if (s_reload) {
            const uint64_t* rw = reinterpret_cast<const uint64_t*>(L.recs + it.slot);
            for (int i = threadIdx.x; i < kRecQwords; i += kBlock) s_rec[i] = sys_load64(rw + i);
        }
        __syncthreads();
        SrvRecord rec;
        uint32_t* dst = reinterpret_cast<uint32_t*>(&rec);
        const uint32_t* src_w = reinterpret_cast<const uint32_t*>(s_rec);
#pragma unroll
        for (int i = 0; i < kRecWords; ++i) dst[i] = __builtin_amdgcn_readfirstlane(src_w[i]);
        if (threadIdx.x == 0) s_strike = rec.strike;
        __syncthreads();

        dispatch<Real, 0, 2>(config_code(rec.d), rec.d, it.tile, &s_strike, L.slots, lds);

        if (threadIdx.x == 0) {  // thread 0 wrote this tile's slot
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            const uint32_t prev =
                __hip_atomic_fetch_add(L.done + it.slot, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            s_last = (prev + 1u == it.target) ? 1u : 0u;
        }
        __syncthreads();
        if (s_last) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            const double2 r = block_sum_slots(L.slots, rec.d.slot_off, it.n_tiles, s_w);
            if (threadIdx.x == 0) {
                const uint64_t idx =
                    __hip_atomic_fetch_add(&L.dc->comp_head, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                SrvCompletion* c = L.comps + (idx % L.comp_cap);
                __hip_atomic_store(&c->slot, it.slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                __hip_atomic_store(&c->target, it.target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                __hip_atomic_store(&c->n_tiles, it.n_tiles, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                __hip_atomic_store(reinterpret_cast<uint64_t*>(&c->sum), __double_as_longlong(r.x), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_SYSTEM);
                __hip_atomic_store(reinterpret_cast<uint64_t*>(&c->sum_sq), __double_as_longlong(r.y),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                __hip_atomic_store(&c->seq, idx + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
            }
        }
        __syncthreads();
    }
}

// One thread: this launch's ticket base is where the previous launch drained;
// its ticket counter and drain flag start at zero.
__global__ void k_server_prepare(SrvMirror* m) {
    m->launch_base = m->next_base;
    m->tickets = 0;
    m->shutdown = 0;
}

}  // namespace

hipError_t server_occupancy(int fp32, int* blocks_per_cu) {
    return fp32 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_server<float>, kBlock, 0)
                : hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_server<double>, kBlock, 0);
}

size_t server_mirror_bytes() { return sizeof(SrvMirror); }

hipError_t launch_server(const SrvLaunch& L, void* mirror, hipStream_t s) {
    if (L.n_blocks < 2) return hipErrorInvalidValue;
    SrvMirror* m = static_cast<SrvMirror*>(mirror);
    hipLaunchKernelGGL(k_server_prepare, dim3(1), dim3(1), 0, s, m);
    if (L.fp32)
        hipLaunchKernelGGL(k_server<float>, dim3(L.n_blocks), dim3(kBlock), 0, s, L, m);
    else
        hipLaunchKernelGGL(k_server<double>, dim3(L.n_blocks), dim3(kBlock), 0, s, L, m);
    return hipGetLastError();
}

}  // namespace ope
