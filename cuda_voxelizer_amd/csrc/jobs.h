// jobs.h -- mapping a launch's blocks to voxelization jobs.
//
// Every kernel serves either one job passed by value (ItemMap.items ==
// nullptr) or a device array of jobs (batched mode); ItemMap.base holds the
// exclusive prefix of the launch's work units (blocks, tiles, triangles) per
// job, so a block finds its job with a binary search.
#pragma once

#include <hip/hip_runtime.h>

#include "tile.h"

namespace voxamd {

struct BlockJob {
    ItemDev it;
    uint32_t local;  // work unit within the job
    int index;       // job index (0 for a single job)
};

// Thread 0 resolves, the block reads the result from LDS.  Units past the
// total (grid rounding) get local = ~0u.
__device__ __forceinline__ void resolve_job(const ItemDev& one, const ItemMap& M, uint32_t unit, BlockJob& J) {
    if (threadIdx.x == 0) {
        if (M.items == nullptr) {
            J.it = one;
            J.local = unit;
            J.index = 0;
        } else if (unit >= M.base[M.nitems]) {
            J.it = M.items[0];
            J.local = ~0u;
            J.index = 0;
        } else {
            int lo = 0, hi = M.nitems - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (M.base[mid] <= unit) lo = mid; else hi = mid - 1;
            }
            J.it = M.items[lo];
            J.local = unit - M.base[lo];
            J.index = lo;
        }
    }
    __syncthreads();
}

// Job of a global triangle id (wide pass): base = per-job triangle prefix.
__device__ __forceinline__ int job_of_triangle(const ItemMap& M, uint32_t gt) {
    if (M.items == nullptr) return 0;
    int lo = 0, hi = M.nitems - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (M.base[mid] <= gt) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ const ItemDev& job_ref(const ItemDev& one, const ItemMap& M, int j) {
    return M.items ? M.items[j] : one;
}

}  // namespace voxamd
