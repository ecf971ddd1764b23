// chunk.hip -- single-pass surface voxelization with LDS word merging.
//
// One workgroup takes a chunk of consecutive triangles (meshes are stored
// with spatially coherent triangle order, so a chunk covers a compact
// region).  The chunk is staged through LDS with coalesced loads, each wave
// rasterizes its 64 triangles with the wave-balanced raster (raster.h), and
// the passing voxels are OR-ed into an LDS hash table keyed by table word, so
// bits that several triangles of the chunk set in one word are merged before
// memory is touched.  The table is then updated with one non-returning global
// atomicOr per distinct word; words that do not fit in the hash table go
// straight to global atomics.  The table is zero-filled beforehand (unless
// accumulating).
#include "jobs.h"
#include "raster.h"
#include "vox_kernels.h"

#pragma clang fp contract(off)

namespace voxamd {

constexpr int kChunk = 256;  // triangles per workgroup (one per lane)
constexpr int kHashBits = 10;
constexpr int kHashCap = 1 << kHashBits;
constexpr uint32_t kEmpty = 0xFFFFFFFFu;

struct ChunkShared {
    uint32_t keys[kHashCap];
    uint32_t vals[kHashCap];
    union {
        float tri[kChunk * 9];     // staging of the chunk's triangles
        WaveRaster wave[kChunk / 64];
    } u;
};

__device__ __forceinline__ void word_or(ChunkShared& S, uint32_t* __restrict__ table, uint32_t key, uint32_t mask) {
    uint32_t h = (key * 2654435761u) >> (32 - kHashBits);
#pragma unroll 1
    for (int probe = 0; probe < 8; probe++) {
        uint32_t k = S.keys[h];
        if (k == kEmpty) {
            k = atomicCAS(&S.keys[h], kEmpty, key);
            if (k == kEmpty) k = key;
        }
        if (k == key) {
            atomicOr(&S.vals[h], mask);
            return;
        }
        h = (h + 1) & (kHashCap - 1);
    }
    atomicOr(table + key, mask);  // hash table crowded: go to memory directly
}

// Work unit: kChunk consecutive triangles of one job.
__global__ void __launch_bounds__(kChunk) surface_chunk_kernel(ItemDev one, ItemMap M) {
    __shared__ ChunkShared S;
    __shared__ BlockJob J;
    resolve_job(one, M, blockIdx.x, J);
    if (J.local == ~0u) return;
    const vox_info info = J.it.info;
    const float* __restrict__ tris = J.it.tris;
    uint32_t* __restrict__ table = J.it.table;
    const int layout = J.it.layout;
    const uint64_t n = info.n_triangles;
    const uint32_t G = info.gridsize[0];
    const uint64_t first = (uint64_t)J.local * kChunk;
    if (first >= n) return;
    const int count = (int)((n - first) < (uint64_t)kChunk ? (n - first) : (uint64_t)kChunk);
    for (int i = threadIdx.x; i < kHashCap; i += kChunk) {
        S.keys[i] = kEmpty;
        S.vals[i] = 0u;
    }
    const float* src = tris + first * 9;
    for (int i = threadIdx.x; i < count * 9; i += kChunk) S.u.tri[i] = src[i];
    __syncthreads();
    const bool have = (int)threadIdx.x < count;
    float v[3][3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        v[k][0] = (have ? S.u.tri[threadIdx.x * 9 + 3 * k + 0] : 0.0f) - info.bbox_min[0];
        v[k][1] = (have ? S.u.tri[threadIdx.x * 9 + 3 * k + 1] : 0.0f) - info.bbox_min[1];
        v[k][2] = (have ? S.u.tri[threadIdx.x * 9 + 3 * k + 2] : 0.0f) - info.bbox_min[2];
    }
    SurfTri T;
    surface_setup(v, info, T);
    __syncthreads();  // staging area is reused by the raster
    WaveRaster& W = S.u.wave[threadIdx.x >> 6];
    if (layout == VOX_LAYOUT_LINEAR && (G & 31u) == 0) {
        // rows are word aligned: one x chunk is one table word
        wave_raster(W, info, have, T, T.lo[0], T.hi[0], T.lo[1], T.hi[1], T.lo[2], T.hi[2],
                    [&](int y, int z, int cx, uint32_t mask) {
                        const uint64_t w = ((uint64_t)cx + (uint64_t)y * G + (uint64_t)z * G * G) >> 5;
                        word_or(S, table, (uint32_t)w, __brev(mask));
                    });
    } else if (layout == VOX_LAYOUT_MORTON) {
        wave_raster(W, info, have, T, T.lo[0], T.hi[0], T.lo[1], T.hi[1], T.lo[2], T.hi[2],
                    [&](int y, int z, int cx, uint32_t mask) {
                        const uint64_t base = morton3((uint32_t)cx, (uint32_t)y, (uint32_t)z);
                        const uint32_t yz_low = ((uint32_t)(y & 1) << 1) | ((uint32_t)(z & 1) << 2) |
                                                ((uint32_t)((y >> 1) & 1) << 4);
                        morton_chunk_words(base, yz_low, mask,
                                           [&](uint64_t w, uint32_t bits) { word_or(S, table, (uint32_t)w, bits); });
                    });
    } else {
        // linear rows not aligned to words (G % 32 != 0): per voxel
        wave_raster(W, info, have, T, T.lo[0], T.hi[0], T.lo[1], T.hi[1], T.lo[2], T.hi[2],
                    [&](int y, int z, int cx, uint32_t mask) {
                        const uint64_t row = (uint64_t)y * G + (uint64_t)z * G * G;
                        uint32_t cur = kEmpty, bits = 0;
                        while (mask) {
                            const int k = __ffs(mask) - 1;
                            mask &= mask - 1;
                            const uint64_t idx = row + (uint64_t)(cx + k);
                            const uint32_t w = (uint32_t)(idx >> 5);
                            if (w != cur) {
                                if (bits) word_or(S, table, cur, bits);
                                cur = w;
                                bits = 0;
                            }
                            bits |= bit_mask(idx);
                        }
                        if (bits) word_or(S, table, cur, bits);
                    });
    }
    for (int i = threadIdx.x; i < kHashCap; i += kChunk) {
        const uint32_t k = S.keys[i];
        if (k != kEmpty) atomicOr(table + k, S.vals[i]);
    }
}

// Zero-fill with 16-byte stores (a kernel, so it can live inside a captured
// graph).  Work unit: kFillWords words of one job's table.
constexpr uint64_t kFillWords = 16384;

__global__ void __launch_bounds__(256) fill_zero_kernel(ItemDev one, ItemMap M) {
    __shared__ BlockJob J;
    resolve_job(one, M, blockIdx.x, J);
    if (J.local == ~0u) return;
    const uint32_t G = J.it.info.gridsize[0];
    const uint64_t words = ((uint64_t)G * G * G + 31) / 32;
    const uint64_t w0 = (uint64_t)J.local * kFillWords;
    const uint64_t w1 = w0 + kFillWords < words ? w0 + kFillWords : words;
    uint32_t* __restrict__ table = J.it.table;
    if ((((uintptr_t)table) & 15u) == 0) {
        uint4* t4 = (uint4*)table;
        const uint64_t a = (w0 + 3) / 4, b = w1 / 4;  // whole uint4 inside [w0, w1)
        for (uint64_t i = a + threadIdx.x; i < b; i += 256) t4[i] = make_uint4(0u, 0u, 0u, 0u);
        for (uint64_t i = w0 + threadIdx.x; i < w1 && i < 4 * a; i += 256) table[i] = 0u;
        for (uint64_t i = (b * 4 > w0 ? b * 4 : w0) + threadIdx.x; i < w1; i += 256) table[i] = 0u;
    } else {
        for (uint64_t i = w0 + threadIdx.x; i < w1; i += 256) table[i] = 0u;
    }
}

bool chunk_supported(uint32_t G) { return (uint64_t)G * G * G / 32 < 0xFFFFFFFFull; }

uint32_t fill_units(uint32_t G) {
    const uint64_t words = ((uint64_t)G * G * G + 31) / 32;
    return (uint32_t)((words + kFillWords - 1) / kFillWords);
}

uint32_t chunk_units(uint64_t n_triangles) { return (uint32_t)((n_triangles + kChunk - 1) / kChunk); }

hipError_t launch_fill_jobs(const ItemDev& one, const ItemMap& M, uint32_t units, hipStream_t stream) {
    if (units == 0) return hipSuccess;
    fill_zero_kernel<<<units, 256, 0, stream>>>(one, M);
    return hipGetLastError();
}

hipError_t launch_chunk_jobs(const ItemDev& one, const ItemMap& M, uint32_t units, hipStream_t stream) {
    if (units == 0) return hipSuccess;
    surface_chunk_kernel<<<units, kChunk, 0, stream>>>(one, M);
    return hipGetLastError();
}

static ItemDev single_job(const vox_info& info, const float* tris, uint32_t* table, int layout) {
    ItemDev one{};
    one.info = info;
    one.tris = tris;
    one.table = table;
    one.layout = layout;
    return one;
}

hipError_t launch_fill_zero(const vox_info& info, uint32_t* table, hipStream_t stream) {
    const ItemMap none{nullptr, nullptr, 1, 0};
    return launch_fill_jobs(single_job(info, nullptr, table, 0), none, fill_units(info.gridsize[0]), stream);
}

hipError_t launch_surface_chunk(const vox_info& info, const float* tris, uint32_t* table, int layout,
                                hipStream_t stream) {
    if (info.n_triangles == 0) return hipSuccess;
    const ItemMap none{nullptr, nullptr, 1, 0};
    return launch_chunk_jobs(single_job(info, tris, table, layout), none, chunk_units(info.n_triangles), stream);
}

}  // namespace voxamd
