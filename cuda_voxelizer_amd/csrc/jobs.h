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

// The block finds its job with one round of parallel loads (thread t tests
// whether unit lies in job t's range; a binary search by one thread would
// wait on ~log2(jobs) dependent loads) and reads the result from LDS.  Units
// past the total (grid rounding) get local = ~0u.
__device__ __forceinline__ void resolve_job(const ItemDev& one, const ItemMap& M, uint32_t unit, BlockJob& J) {
    if (M.items == nullptr) {
        if (threadIdx.x == 0) {
            J.it = one;
            J.local = unit;
            J.index = 0;
        }
    } else {
        if (threadIdx.x == 0) {
            J.local = ~0u;  // unless a job claims the unit below
            J.index = 0;
        }
        __syncthreads();
        for (int j0 = 0; j0 < M.nitems; j0 += (int)blockDim.x) {
            const int j = j0 + (int)threadIdx.x;
            if (j < M.nitems) {
                const uint32_t lo = M.base[j], hi = M.base[j + 1];
                if (lo <= unit && unit < hi) {  // at most one job (empty ones never match)
                    J.it = M.items[j];
                    J.local = unit - lo;
                    J.index = j;
                }
            }
        }
        __syncthreads();
        if (threadIdx.x == 0 && J.local == ~0u) J.it = M.items[0];
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
