// lp_queue.h -- persistent launches over a (frame, tile) work queue.
//
// The grid form of a per-point kernel launches (capacity tiles) x (frames)
// workgroups; tiles beyond a frame's point count exit at once.  The
// persistent form launches a fixed number of workgroups that take (frame,
// tile) items from per-XCD queues until they run dry: the items of the
// frames f with f % Q == q (Q = min(8, frames)) form queue q, a frame's
// tiles in order, and a workgroup with linear id w (the dispatcher hands
// workgroups to the 8 XCDs round-robin) takes items from queue (w % 8) % Q,
// so one frame's tiles stay in one XCD's L2.  Each queue is one counter,
// advanced with an atomic by thread 0 of the workgroup.  The counters are
// cleared by the per-run clear kernel, ahead of every persistent launch.
#pragma once

#include "lp_device.h"
#include "lp_internal.h"

namespace lp {

// Queues of one persistent launch site (8 counters each).
enum WqSite : int { kWqProject = 0, kWqSegLabels, kWqSpherical, kWqVoxelInsert, kWqSites };
static_assert(kWqCounters == kWqSites * 8, "one counter per XCD queue and site");
constexpr int kPersistentWorkgroups = 1024;  // 256 CUs x 4

// Runs body(frame, tile) for every tile of every frame of the batch, as
// items of the per-XCD queues.  count(f) is frame f's item count in units
// of `tile_size` elements.  All threads of the workgroup call it.
template <class CountFn, class Body>
__device__ __forceinline__ void wq_run(uint32_t* counters, int B, int tile_size, CountFn count, Body body) {
    __shared__ uint32_t item_sh;
    const int Q = B < 8 ? B : 8;
    const int q = static_cast<int>(blockIdx.x % 8u) % Q;
    uint32_t* ctr = counters + q;
    while (true) {
        if (threadIdx.x == 0) item_sh = atomicAdd(ctr, 1u);
        __syncthreads();
        uint32_t item = item_sh;
        __syncthreads();
        int f = -1, tile = 0;
        for (int g = q; g < B; g += Q) {
            const uint32_t n = static_cast<uint32_t>(count(g));
            const uint32_t tiles = (n + tile_size - 1) / tile_size;
            if (item < tiles) {
                f = g;
                tile = static_cast<int>(item);
                break;
            }
            item -= tiles;
        }
        if (f < 0) return;  // queue exhausted: uniform across the workgroup
        body(f, tile);
    }
}

}  // namespace lp
