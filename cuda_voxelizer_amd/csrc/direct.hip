// direct.hip -- one-triangle-per-lane kernels with global atomics.
//
// This is the plain launch shape of the reference (voxelize.cu:58-190 and
// voxelize_solid.cu:73-145: one thread per triangle, an atomic per written
// word) kept as a comparison path and as the fallback for tables too large
// for the tile binning workspace.  Differences from the reference kernels:
// the arithmetic follows the CPU path bit for bit (vox_device.h), solid runs
// XOR whole words instead of single bits, and bits of one word found by one
// lane are merged before the atomic.
#include "jobs.h"
#include "vox_device.h"
#include "vox_kernels.h"

namespace voxamd {

// Work unit: 256 consecutive triangles of one job, one per lane.
__global__ void __launch_bounds__(256) surface_direct_kernel(ItemDev one, ItemMap M) {
    __shared__ BlockJob J;
    resolve_job(one, M, blockIdx.x, J);
    if (J.local == ~0u) return;
    const vox_info info = J.it.info;
    const float* __restrict__ tris = J.it.tris;
    uint32_t* __restrict__ table = J.it.table;
    const int layout = J.it.layout;
    const uint64_t n = info.n_triangles;
    const uint32_t G = info.gridsize[0];
    {
        const uint64_t t = (uint64_t)J.local * 256 + threadIdx.x;
        if (t >= n) return;
        float v[3][3];
        load_tri(tris, t, info, v);
        SurfTri T;
        surface_setup(v, info, T);
        for (int z = T.lo[2]; z <= T.hi[2]; z++) {
            const float pz = (float)z * info.unit[2];
            for (int y = T.lo[1]; y <= T.hi[1]; y++) {
                const float py = (float)y * info.unit[1];
                if (!edges_pass(T.yz, py, pz)) continue;
                uint64_t cur_word = ~0ull;
                uint32_t cur_mask = 0;
                for (int x = T.lo[0]; x <= T.hi[0]; x++) {
                    const float px = (float)x * info.unit[0];
                    if (!(plane_pass(T, px, py, pz) && edges_pass(T.xy, px, py) && edges_pass(T.zx, pz, px)))
                        continue;
                    const uint64_t idx = voxel_index(layout, G, x, y, z);
                    const uint64_t w = idx >> 5;
                    if (w != cur_word) {
                        if (cur_mask) atomicOr(&table[cur_word], cur_mask);
                        cur_word = w;
                        cur_mask = 0;
                    }
                    cur_mask |= bit_mask(idx);
                }
                if (cur_mask) atomicOr(&table[cur_word], cur_mask);
            }
        }
    }
}

__global__ void __launch_bounds__(256) solid_direct_kernel(ItemDev one, ItemMap M) {
    __shared__ BlockJob J;
    resolve_job(one, M, blockIdx.x, J);
    if (J.local == ~0u) return;
    const vox_info info = J.it.info;
    const float* __restrict__ tris = J.it.tris;
    uint32_t* __restrict__ table = J.it.table;
    const int layout = J.it.layout;
    unsigned long long* __restrict__ offrow_out = J.it.offrow;
    const uint64_t n = info.n_triangles;
    const int G = (int)info.gridsize[0];
    const uint64_t total_bits = (uint64_t)G * (uint64_t)G * (uint64_t)G;
    unsigned long long offrow = 0;
    {
        const uint64_t t = (uint64_t)J.local * 256 + threadIdx.x;
        if (t >= n) return;
        float v[3][3];
        load_tri(tris, t, info, v);
        SolidTri S;
        solid_setup(v, info, S);
        if (!S.valid) return;
        int ylo = (int)S.ymin, zlo = (int)S.zmin;
        if (ylo < 0) { offrow++; ylo = 0; }
        if (zlo < 0) { offrow++; zlo = 0; }
        if (S.ymax > (float)(G - 1)) offrow++;
        if (S.zmax > (float)(G - 1)) offrow++;
        for (int y = ylo; (float)y <= S.ymax && y < G; y++) {
            const float py = ((float)y + 0.5f) * info.unit[1];
            for (int z = zlo; (float)z <= S.zmax && z < G; z++) {
                const float pz = ((float)z + 0.5f) * info.unit[2];
                if (!column_hit(S, py, pz)) continue;
                int64_t xmax = column_xmax(S, info, py, pz);
                if (xmax < 0 || xmax >= G) offrow++;
                if (xmax < 0) continue;
                if (layout == VOX_LAYOUT_LINEAR) {
                    const uint64_t b0 = (uint64_t)y * (uint64_t)G + (uint64_t)z * (uint64_t)G * (uint64_t)G;
                    const uint64_t room = total_bits - 1 - b0;  // last bit index inside the table
                    const uint64_t b1 = b0 + ((uint64_t)xmax < room ? (uint64_t)xmax : room);  // inclusive
                    for (uint64_t w = b0 >> 5; w <= (b1 >> 5); w++) {
                        const uint32_t s = (w == (b0 >> 5)) ? (uint32_t)(b0 & 31u) : 0u;
                        const uint32_t e = (w == (b1 >> 5)) ? (uint32_t)(b1 & 31u) : 31u;
                        const uint32_t mask = (0xFFFFFFFFu >> s) & (0xFFFFFFFFu << (31u - e));
                        atomicXor(&table[w], mask);
                    }
                } else {
                    const int xe = xmax < G ? (int)xmax : G - 1;  // Morton codes of x >= G are past the table
                    uint64_t cur_word = ~0ull;
                    uint32_t cur_mask = 0;
                    for (int x = 0; x <= xe; x++) {
                        const uint64_t idx = morton3((uint32_t)x, (uint32_t)y, (uint32_t)z);
                        if ((idx >> 5) != cur_word) {
                            if (cur_mask) atomicXor(&table[cur_word], cur_mask);
                            cur_word = idx >> 5;
                            cur_mask = 0;
                        }
                        cur_mask |= bit_mask(idx);
                    }
                    if (cur_mask) atomicXor(&table[cur_word], cur_mask);
                }
            }
        }
    }
    if (offrow_out && offrow) atomicAdd(offrow_out, offrow);
}

uint32_t direct_units(uint64_t n_triangles) { return (uint32_t)((n_triangles + 255) / 256); }

hipError_t launch_direct_jobs(const ItemDev& one, const ItemMap& M, uint32_t units, bool solid, hipStream_t stream) {
    if (units == 0) return hipSuccess;
    if (solid)
        solid_direct_kernel<<<units, 256, 0, stream>>>(one, M);
    else
        surface_direct_kernel<<<units, 256, 0, stream>>>(one, M);
    return hipGetLastError();
}

hipError_t launch_surface_direct(const vox_info& info, const float* tris, uint32_t* table, int layout,
                                 hipStream_t stream) {
    if (info.n_triangles == 0) return hipSuccess;
    ItemDev one{};
    one.info = info;
    one.tris = tris;
    one.table = table;
    one.layout = layout;
    const ItemMap none{nullptr, nullptr, 1, 0};
    return launch_direct_jobs(one, none, direct_units(info.n_triangles), false, stream);
}

hipError_t launch_solid_direct(const vox_info& info, const float* tris, uint32_t* table, int layout,
                               unsigned long long* offrow, hipStream_t stream) {
    if (info.n_triangles == 0) return hipSuccess;
    ItemDev one{};
    one.info = info;
    one.tris = tris;
    one.table = table;
    one.layout = layout;
    one.offrow = offrow;
    one.solid = 1;
    const ItemMap none{nullptr, nullptr, 1, 0};
    return launch_direct_jobs(one, none, direct_units(info.n_triangles), true, stream);
}

}  // namespace voxamd
