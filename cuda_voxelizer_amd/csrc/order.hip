// order.hip -- vox_sort_triangles: a mesh's triangles in spatial order.
//
// The chunk kernel merges the words of 256 consecutive triangles in an LDS
// hash and ORs each distinct word into memory with one memory-side atomic; a
// word that triangles of several chunks set costs one atomic per chunk.  In
// the order a mesh is generated or stored, neighbouring triangles can sit far
// apart: the 1024^3 torus (1,048,576 triangles, rings of 1,024) has 867,526
// distinct words but its chunks flush 1.85 M, and the faces of a subdivided
// icosphere whose levels interleave its children are scattered over the whole
// sphere.  Bucketing the triangles by the 16^3-voxel brick of their voxel box's
// lowest corner, bricks in Morton order, puts each brick's triangles into the
// same or neighbouring chunks.  Done once for a mesh that is voxelized many
// times (the sort costs about as much as it saves in one call).
//
// A counting sort in four launches:
//   keys:    one key per triangle, and per-bucket counts (atomics, one per
//            run of equal keys in a wave);
//   totals:  one workgroup per group of 1024 buckets: the group's total;
//   scan:    one workgroup per group: exclusive prefix of its buckets, offset
//            by the totals of the groups before it; the counts are returned
//            to zero for the next call;
//   scatter: each run of equal keys in a wave reserves its slots with one
//            atomicAdd on its bucket's cursor, and each lane copies its
//            triangle (and its index, if asked) into its slot.
// Group totals counted with atomics in the keys pass serialised on their few
// addresses (1.6 ms for 1.3 M triangles in random order), hence the totals pass.
// The order within a bucket follows the order in which the waves'
// reservations land.  Bricks are 16 voxels on a side, 2^shift for grids
// above 1024 so that each axis has at most 64: at most 2^18 buckets.

#include "vox_kernels.h"
#include "vox_device.h"
#include "vox_host.h"
#include "raster.h"
#include "scan.h"
#include <map>
#include <mutex>

namespace voxamd {

constexpr int kOrderGroup = 1024;  // buckets per scan workgroup (256 threads x 4)

struct OrderGeom {
    int shift;  // brick edge 2^shift voxels
    uint32_t buckets;
    uint32_t groups;
};

static OrderGeom order_geom(uint32_t G) {
    int lg = 0;
    while ((1u << lg) < G) lg++;
    OrderGeom g;
    g.shift = lg - 6 > 4 ? lg - 6 : 4;
    const uint32_t per_axis = (G + (1u << g.shift) - 1) >> g.shift;  // at most 64
    int ab = 1;  // at least 8 buckets: the scan moves them in groups of 4
    while ((1u << ab) < per_axis) ab++;
    g.buckets = 1u << (3 * ab);
    g.groups = (g.buckets + kOrderGroup - 1) / kOrderGroup;
    return g;
}

// Lanes holding a triangle are a prefix of the wave (t = first + lane < n).
// For each run of equal keys among them: whether this lane starts one, the
// run's length (at its first lane) and the run's first lane (at every lane).
__device__ __forceinline__ void wave_runs(uint32_t key, bool live, bool& start, int& len, int& head) {
    const int lane = (int)__lane_id();
    const uint32_t prev = __shfl_up(key, 1);
    start = live && (lane == 0 || prev != key);
    const unsigned long long starts = __ballot(start);
    const int nlive = (int)__popcll(__ballot(live));
    const unsigned long long upto = (2ull << lane) - 1ull;  // lanes 0..lane (all lanes when lane == 63)
    const unsigned long long after = starts & ~upto;
    len = (after ? (int)__builtin_ctzll(after) : nlive) - lane;
    const unsigned long long before = starts & upto;
    head = before ? 63 - (int)__builtin_clzll(before) : 0;
}

__global__ void __launch_bounds__(256) order_keys_kernel(const float* __restrict__ tris, vox_info info, uint64_t n,
                                                         int shift, uint32_t key_mask, uint32_t* __restrict__ keys,
                                                         uint32_t* __restrict__ counts) {
    const uint64_t t = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    const bool live = t < n;
    uint32_t key = 0u;
    if (live) {
        float v[3][3];
        load_tri(tris, t, info, v);
        int lo[3], hi[3];
        surface_box(v, info, lo, hi);  // clamped to [0, G-1]
        key = spread10((uint32_t)lo[0] >> shift) | (spread10((uint32_t)lo[1] >> shift) << 1) |
              (spread10((uint32_t)lo[2] >> shift) << 2);
        // below the bucket count already (the box is clamped to the grid); the
        // mask keeps the index in range whatever a NaN vertex made of the box
        key &= key_mask;
        keys[t] = key;
    }
    bool start;
    int len, head;
    wave_runs(key, live, start, len, head);
    if (start) atomicAdd(counts + key, (uint32_t)len);
}

// One workgroup per group of kOrderGroup buckets (buckets is a multiple of 4).
__global__ void __launch_bounds__(256) order_totals_kernel(const uint32_t* __restrict__ counts,
                                                           uint32_t* __restrict__ group_totals, uint32_t buckets) {
    const uint32_t i0 = blockIdx.x * kOrderGroup + threadIdx.x * 4u;
    uint4 c = make_uint4(0u, 0u, 0u, 0u);
    if (i0 < buckets) c = *(const uint4*)(counts + i0);
    uint64_t total;
    (void)block_excl_scan<SumOp>((uint64_t)c.x + c.y + c.z + c.w, &total);
    if (threadIdx.x == 0) group_totals[blockIdx.x] = (uint32_t)total;
}

__global__ void __launch_bounds__(256) order_scan_kernel(uint32_t* __restrict__ counts,
                                                         uint32_t* __restrict__ cursor,
                                                         const uint32_t* __restrict__ group_totals,
                                                         uint32_t buckets) {
    const uint32_t g = blockIdx.x;
    // offset of this group: the totals of the groups before it (at most 256)
    uint64_t total;
    (void)block_excl_scan<SumOp>(threadIdx.x < g ? group_totals[threadIdx.x] : 0u, &total);
    const uint32_t base = (uint32_t)total;
    const uint32_t i0 = g * kOrderGroup + threadIdx.x * 4u;
    uint4 c = make_uint4(0u, 0u, 0u, 0u);
    if (i0 < buckets) c = *(const uint4*)(counts + i0);
    const uint64_t excl = block_excl_scan<SumOp>((uint64_t)c.x + c.y + c.z + c.w, &total);
    if (i0 < buckets) {
        const uint32_t a = base + (uint32_t)excl;
        *(uint4*)(cursor + i0) = make_uint4(a, a + c.x, a + c.x + c.y, a + c.x + c.y + c.z);
        *(uint4*)(counts + i0) = make_uint4(0u, 0u, 0u, 0u);  // zero for the next call
    }
}

__global__ void __launch_bounds__(256) order_scatter_kernel(const float* __restrict__ tris,
                                                            const uint32_t* __restrict__ keys, uint64_t n,
                                                            uint32_t* __restrict__ cursor, float* __restrict__ out,
                                                            uint32_t* __restrict__ perm) {
    const uint64_t t = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    const bool live = t < n;
    const uint32_t key = live ? keys[t] : 0u;
    bool start;
    int len, head;
    wave_runs(key, live, start, len, head);
    uint32_t base = 0u;
    if (start) base = atomicAdd(cursor + key, (uint32_t)len);
    base = __shfl(base, head);
    const uint64_t slot = (uint64_t)base + (uint64_t)((int)__lane_id() - head);
    if (live && slot < n) {  // always below n while the counts start at zero
        const float* __restrict__ src = tris + 9 * t;
        float* __restrict__ dst = out + 9 * slot;
#pragma unroll
        for (int k = 0; k < 9; k++) dst[k] = src[k];
        if (perm) perm[slot] = (uint32_t)t;
    }
}

__global__ void __launch_bounds__(256) order_zero_kernel(uint32_t* __restrict__ p, uint64_t words) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < words; i += (uint64_t)gridDim.x * 256u)
        p[i] = 0u;
}

// ---- per (device, stream) buffers ------------------------------------------
// [group totals: 256][counts: buckets][cursor: buckets][keys: n], uint32
// each.  The counts are zero between calls (the scan kernel returns them to
// zero); `dirty` marks buffers whose zeroing is not known to have been queued.
struct OrderBuffers {
    void* ptr = nullptr;
    uint64_t n_cap = 0;
    uint32_t buckets_cap = 0;
    bool dirty = true;
};
static std::mutex g_order_mutex;
static std::map<std::pair<int, hipStream_t>, OrderBuffers> g_order;

static hipError_t sort_triangles(const vox_info& info, const float* tris, float* out, uint32_t* perm,
                                 hipStream_t stream) {
    const uint64_t n = info.n_triangles;
    const OrderGeom g = order_geom(info.gridsize[0]);
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    uint32_t *group_totals, *counts, *cursor, *keys;
    {
        std::lock_guard<std::mutex> lock(g_order_mutex);
        OrderBuffers& b = g_order[{dev, stream}];
        if (b.ptr == nullptr || n > b.n_cap || g.buckets > b.buckets_cap) {
            const uint64_t n_cap = n > b.n_cap ? n : b.n_cap;
            const uint32_t buckets_cap = g.buckets > b.buckets_cap ? g.buckets : b.buckets_cap;
            if (b.ptr) {
                e = hipStreamSynchronize(stream);  // a previous call on this stream may still use them
                if (e == hipSuccess) e = hipFree(b.ptr);
                b = OrderBuffers{};
                if (e != hipSuccess) return e;
            }
            e = hipMalloc(&b.ptr, 4ull * (256u + 2ull * buckets_cap + n_cap));
            if (e != hipSuccess) {
                b = OrderBuffers{};
                return e;
            }
            b.n_cap = n_cap;
            b.buckets_cap = buckets_cap;
            b.dirty = true;
        }
        group_totals = (uint32_t*)b.ptr;
        counts = group_totals + 256;
        cursor = counts + b.buckets_cap;
        keys = cursor + b.buckets_cap;
        if (b.dirty) {
            const uint64_t blocks = (b.buckets_cap + 255) / 256;
            order_zero_kernel<<<(unsigned)(blocks < 1024 ? blocks : 1024), 256, 0, stream>>>(counts, b.buckets_cap);
        }
        b.dirty = true;  // until the scan kernel is queued
    }
    const unsigned blocks = (unsigned)((n + 255) / 256);
    order_keys_kernel<<<blocks, 256, 0, stream>>>(tris, info, n, g.shift, g.buckets - 1u, keys, counts);
    order_totals_kernel<<<g.groups, 256, 0, stream>>>(counts, group_totals, g.buckets);
    order_scan_kernel<<<g.groups, 256, 0, stream>>>(counts, cursor, group_totals, g.buckets);
    order_scatter_kernel<<<blocks, 256, 0, stream>>>(tris, keys, n, cursor, out, perm);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lock(g_order_mutex);
    g_order[{dev, stream}].dirty = false;
    return hipSuccess;
}

}  // namespace voxamd

using namespace voxamd;

extern "C" vox_status vox_sort_triangles(const vox_info* info, const float* d_tris, float* d_out, uint32_t* d_perm,
                                         void* stream) {
    if (!info) return fail(VOX_ERR_INVALID, "info is NULL");
    if (info->n_triangles == 0) return VOX_OK;  // nothing to order (the buffers may be NULL)
    // the voxel boxes are taken at info's grid: the same checks as a voxelization
    // (d_out stands in for the table pointer)
    vox_status s = validate(info, d_tris, (const uint32_t*)d_out, VOX_LAYOUT_LINEAR);
    if (s != VOX_OK) return s;
    const uint64_t n = info->n_triangles;
    const uintptr_t a0 = (uintptr_t)d_tris, a1 = a0 + 36ull * n, b0 = (uintptr_t)d_out, b1 = b0 + 36ull * n;
    if (a0 < b1 && b0 < a1) return fail(VOX_ERR_INVALID, "d_out overlaps d_tris (the sort is out of place)");
    if (d_perm) {
        const uintptr_t p0 = (uintptr_t)d_perm, p1 = p0 + 4ull * n;
        if ((p0 < a1 && a0 < p1) || (p0 < b1 && b0 < p1))
            return fail(VOX_ERR_INVALID, "d_perm overlaps d_tris or d_out");
    }
    hipError_t e = sort_triangles(*info, d_tris, d_out, d_perm, (hipStream_t)stream);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        return fail(VOX_ERR_OOM, "triangle order buffers");
    }
    return e == hipSuccess ? VOX_OK : hip_fail(e, "triangle order kernels");
}
