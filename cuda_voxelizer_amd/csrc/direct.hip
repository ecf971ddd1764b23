// direct.hip -- one-triangle-per-lane kernels with global atomics.
//
// This is the plain launch shape of the reference (voxelize.cu:58-190 and
// voxelize_solid.cu:73-145: one thread per triangle, an atomic per written
// word) kept as a comparison path and as the fallback for tables too large
// for the tile binning workspace.  Differences from the reference kernels:
// the arithmetic follows the CPU path bit for bit (vox_device.h), solid runs
// XOR whole words instead of single bits, and bits of one word found by one
// lane are merged before the atomic.
#include "vox_device.h"
#include "vox_kernels.h"

namespace voxamd {

__global__ void __launch_bounds__(256) surface_direct_kernel(vox_info info, const float* __restrict__ tris,
                                                             uint32_t* __restrict__ table, int layout) {
    const uint64_t n = info.n_triangles;
    const uint32_t G = info.gridsize[0];
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < n;
         t += (uint64_t)gridDim.x * blockDim.x) {
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

__global__ void __launch_bounds__(256) solid_direct_kernel(vox_info info, const float* __restrict__ tris,
                                                           uint32_t* __restrict__ table, int layout,
                                                           unsigned long long* __restrict__ offrow_out) {
    const uint64_t n = info.n_triangles;
    const int G = (int)info.gridsize[0];
    const uint64_t total_bits = (uint64_t)G * (uint64_t)G * (uint64_t)G;
    unsigned long long offrow = 0;
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < n;
         t += (uint64_t)gridDim.x * blockDim.x) {
        float v[3][3];
        load_tri(tris, t, info, v);
        SolidTri S;
        solid_setup(v, info, S);
        if (!S.valid) continue;
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

hipError_t launch_surface_direct(const vox_info& info, const float* tris, uint32_t* table, int layout,
                                 hipStream_t stream) {
    if (info.n_triangles == 0) return hipSuccess;
    const uint64_t blocks = (info.n_triangles + 255) / 256;
    const unsigned grid = (unsigned)(blocks < 65535ull * 64 ? blocks : 65535ull * 64);
    surface_direct_kernel<<<grid, 256, 0, stream>>>(info, tris, table, layout);
    return hipGetLastError();
}

hipError_t launch_solid_direct(const vox_info& info, const float* tris, uint32_t* table, int layout,
                               unsigned long long* offrow, hipStream_t stream) {
    if (info.n_triangles == 0) return hipSuccess;
    const uint64_t blocks = (info.n_triangles + 255) / 256;
    const unsigned grid = (unsigned)(blocks < 65535ull * 64 ? blocks : 65535ull * 64);
    solid_direct_kernel<<<grid, 256, 0, stream>>>(info, tris, table, layout, offrow);
    return hipGetLastError();
}

}  // namespace voxamd
