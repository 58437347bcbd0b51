// Shared layout of the persistent serving loop (host: server.cpp, device:
// mc_server.hip).
//
// Host -> device (pinned, fine-grained host memory; the device only loads):
//   SrvHostCtrl.tail      number of work items published (release store by the host)
//   SrvHostCtrl.shutdown  non-zero: the kernel drains (published items still run)
//   SrvItem ring[]        work item i at ring[i % ring_cap]
//   SrvRecord recs[]      descriptor + strike of each resident request slot
// Device -> host (pinned; the device only stores, the host only loads):
//   SrvCompletion comps[] round-completion records; seq = index + 1 is stored
//                         last with release semantics
// Device only (device memory; device-scope atomics):
//   SrvDevCtrl.comp_head  completion records issued
//   done[slot]            tiles finished for this slot, cumulative over every
//                         request that ever used the slot (never reset)
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "mc_desc.h"

namespace ope {

struct SrvHostCtrl {
    uint64_t tail;
    uint32_t shutdown;
    uint32_t pad_;
};

struct SrvItem {
    uint32_t slot;     // request slot
    uint32_t tile;     // tile within the request (kBlock pairs per tile)
    uint32_t target;   // done[slot] value at which this item's round is complete
    uint32_t n_tiles;  // tiles of the request posted so far (reduced at round end)
};

struct alignas(16) SrvRecord {
    McDesc d;
    double strike;
    double pad_;
};

struct alignas(16) SrvCompletion {
    uint64_t seq;  // index + 1 once the record is complete
    uint32_t slot;
    uint32_t target;
    uint32_t n_tiles;
    uint32_t pad_;
    double sum;
    double sum_sq;
};

struct SrvDevCtrl {
    uint64_t comp_head;
    uint64_t pad_;
};

struct SrvLaunch {
    SrvHostCtrl* hc;
    const SrvItem* ring;
    uint32_t ring_cap;
    const SrvRecord* recs;
    SrvDevCtrl* dc;
    uint32_t* done;
    double2* slots;
    SrvCompletion* comps;
    uint32_t comp_cap;
    uint32_t idle_limit;  // polls without a new tail or ticket before the sentinel drains the kernel
    uint32_t n_blocks;    // workgroup 0 is the sentinel, 1.. are workers
    int fp32;
};

static_assert(sizeof(SrvItem) == 16, "work item layout");

// mc_server.hip.  launch_server first resets the launch words of the mirror
// (zero-initialised device memory of server_mirror_bytes() bytes) on the same
// stream, then launches the persistent kernel.
hipError_t launch_server(const SrvLaunch& L, void* mirror, hipStream_t s);
hipError_t server_occupancy(int fp32, int* blocks_per_cu);
size_t server_mirror_bytes();
static_assert(sizeof(SrvRecord) % 4 == 0, "record copied as 32-bit words");

}  // namespace ope
