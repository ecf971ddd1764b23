// chunk.hip -- single-pass surface voxelization with LDS word merging.
//
// One workgroup takes a chunk of 256 consecutive triangles (meshes are stored
// with spatially coherent triangle order, so a chunk covers a compact
// region).  The chunk is staged through LDS with coalesced loads; each lane
// rasterizes its own triangle (ZX edges once per z slice, then per row the YZ
// edges and, at the x that passed ZX, the plane and XY edges) and ORs the passing
// bits into an LDS hash table keyed by table word, so bits that several
// triangles of the chunk set in one word are merged before memory is
// touched.  Triangles with a large voxel box are queued and their rows split
// over the whole workgroup afterwards.  The table is then updated with one
// non-returning global atomicOr per distinct word; words that do not fit in
// the hash go straight to global atomics.  The table is zero-filled
// beforehand (unless accumulating).
#include "jobs.h"
#include "raster.h"
#include "vox_kernels.h"

#pragma clang fp contract(off)

namespace voxamd {

constexpr int kChunk = 256;  // triangles per workgroup (one per lane)
constexpr int kHashBits = 11;
constexpr int kHashCap = 1 << kHashBits;
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr int kQueue = kChunk;  // big triangles split over the workgroup (room for all of them)
constexpr int kBigVox = 256;    // voxel boxes above this (or wider than 32 in x) are queued

struct ChunkShared {
    uint32_t keys[kHashCap];
    uint32_t vals[kHashCap];
    float tri[kChunk * 9];
    unsigned queue[kQueue];
    unsigned qn;
};

__device__ __forceinline__ void word_or(ChunkShared& S, uint32_t* __restrict__ table, uint32_t key, uint32_t mask) {
    uint32_t h = (key * 2654435761u) >> (32 - kHashBits);
#pragma unroll 1
    for (int probe = 0; probe < 16; probe++) {
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

// Passing voxels of row (y, z) as a mask (bit k = voxel cx + k, cx a multiple
// of 32) -> hash, in the table's layout.
__device__ __forceinline__ void emit_mask(ChunkShared& S, uint32_t* __restrict__ table, const vox_info& info,
                                          int layout, int cx, int y, int z, uint32_t mask) {
    if (!mask) return;
    const uint32_t G = info.gridsize[0];
    if (layout == VOX_LAYOUT_MORTON) {
        // codes of grids up to 1024 fit 30 bits: 32-bit spreads
        const uint64_t base = G <= 1024u ? (uint64_t)(spread10((uint32_t)cx) | (spread10((uint32_t)y) << 1) |
                                                      (spread10((uint32_t)z) << 2))
                                         : morton3((uint32_t)cx, (uint32_t)y, (uint32_t)z);
        const uint32_t yz_low = ((uint32_t)(y & 1) << 1) | ((uint32_t)(z & 1) << 2) | ((uint32_t)((y >> 1) & 1) << 4);
        morton_chunk_words(base, yz_low, mask, [&](uint64_t w, uint32_t bits) { word_or(S, table, (uint32_t)w, bits); });
    } else if ((G & 31u) == 0) {
        const uint64_t w = ((uint64_t)cx + (uint64_t)y * G + (uint64_t)z * G * G) >> 5;
        word_or(S, table, (uint32_t)w, __brev(mask));
    } else {
        // linear rows not aligned to words: per voxel
        const uint64_t row = (uint64_t)y * G + (uint64_t)z * G * G;
        uint32_t m = mask, cur = kEmpty, bits = 0;
        while (m) {
            const int k = __ffs(m) - 1;
            m &= m - 1;
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
    }
}

// Row (y, z) voxels x0..x1 inside the 32-voxel chunk at cx -> hash.
__device__ __forceinline__ void chunk_row(ChunkShared& S, uint32_t* __restrict__ table, const SurfTri& T,
                                          const float* xy, const float* zx, const vox_info& info, int layout,
                                          int x0, int x1, int cx, int y, int z, float py, float pz) {
    emit_mask(S, table, info, layout, cx, y, z, row_mask(T.n, T.d1, T.d2, xy, zx, info.unit[0], py, pz, x0, x1, cx));
}

// A lane's own box (x extent <= 32): the ZX edge tests depend on (z, x) only,
// so they are evaluated once per z slice into a mask over the box's x range;
// a row then tests the XY edges and the plane only at the x that passed ZX.
// Every test is the same binary32 expression as in vox_device.h, evaluated
// for a subset of the voxels: the resulting bits are unchanged.
__device__ __forceinline__ void lane_raster(ChunkShared& S, uint32_t* __restrict__ table, const SurfTri& T,
                                            const vox_info& info, int layout) {
    const float ux = info.unit[0], uy = info.unit[1], uz = info.unit[2];
    const int xa = T.lo[0], xb = T.hi[0];
    const int c0 = xa & ~31, sh = xa - c0;
    for (int z = T.lo[2]; z <= T.hi[2]; z++) {
        const float pz = (float)z * uz;
        float azx[3];
#pragma unroll
        for (int i = 0; i < 3; i++) azx[i] = T.zx.a[i] * pz;
        uint32_t zxm = 0;  // bit i: x = xa + i passes the ZX edges
        for (int x = xa; x <= xb; x++) {
            const float px = (float)x * ux;
            bool ok = true;
#pragma unroll
            for (int i = 0; i < 3; i++) {
                float b = azx[i];
                b = b + T.zx.b[i] * px;
                b = b + T.zx.d[i];
                ok = ok && !(b < 0.0f);
            }
            if (ok) zxm |= 1u << (x - xa);
        }
        if (!zxm) continue;
        const float nzpz = T.n[2] * pz;
        for (int y = T.lo[1]; y <= T.hi[1]; y++) {
            const float py = (float)y * uy;
            if (!edges_pass(T.yz, py, pz)) continue;
            const float nypy = T.n[1] * py;
            float bxy[3];
#pragma unroll
            for (int i = 0; i < 3; i++) bxy[i] = T.xy.b[i] * py;
            uint32_t m = 0, cand = zxm;
            while (cand) {
                const int k = __ffs(cand) - 1;
                cand &= cand - 1;
                const float px = (float)(xa + k) * ux;
                float t = T.n[0] * px;
                t = t + nypy;
                t = t + nzpz;
                const float r = (t + T.d1) * (t + T.d2);
                bool ok = !(r > 0.0f);
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    float a = T.xy.a[i] * px;
                    a = a + bxy[i];
                    a = a + T.xy.d[i];
                    ok = ok && !(a < 0.0f);
                }
                if (ok) m |= 1u << k;
            }
            if (!m) continue;
            const uint64_t M = (uint64_t)m << sh;  // bit k = voxel c0 + k
            emit_mask(S, table, info, layout, c0, y, z, (uint32_t)M);
            if (M >> 32) emit_mask(S, table, info, layout, c0 + 32, y, z, (uint32_t)(M >> 32));
        }
    }
}

__device__ __forceinline__ void tri_coeffs(const SurfTri& T, float xy[9], float zx[9]) {
#pragma unroll
    for (int i = 0; i < 3; i++) {
        xy[i] = T.xy.a[i]; xy[3 + i] = T.xy.b[i]; xy[6 + i] = T.xy.d[i];
        zx[i] = T.zx.a[i]; zx[3 + i] = T.zx.b[i]; zx[6 + i] = T.zx.d[i];
    }
}

__device__ __forceinline__ void staged_vertices(const ChunkShared& S, int lt, const vox_info& info, float v[3][3]) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
        v[k][0] = S.tri[lt * 9 + 3 * k + 0] - info.bbox_min[0];
        v[k][1] = S.tri[lt * 9 + 3 * k + 1] - info.bbox_min[1];
        v[k][2] = S.tri[lt * 9 + 3 * k + 2] - info.bbox_min[2];
    }
}

// Work unit: kChunk consecutive triangles of one job.  MULTI: the job comes
// from the launch's job map; otherwise it is `one` (kernel arguments).
template <bool MULTI>
__global__ void __launch_bounds__(kChunk) surface_chunk_kernel(ItemDev one, ItemMap M) {
    __shared__ ChunkShared S;
    __shared__ BlockJob J;
    uint32_t local = blockIdx.x;
    if (MULTI) {
        resolve_job(one, M, blockIdx.x, J);
        local = J.local;
        if (local == ~0u) return;
    }
    const ItemDev& it = MULTI ? J.it : one;
    const vox_info& info = it.info;
    uint32_t* __restrict__ table = it.table;
    const int layout = it.layout;
    const uint64_t n = info.n_triangles;
    const uint64_t first = (uint64_t)local * kChunk;
    if (first >= n) return;
    const int count = (int)((n - first) < (uint64_t)kChunk ? (n - first) : (uint64_t)kChunk);
    for (int i = threadIdx.x; i < kHashCap; i += kChunk) {
        S.keys[i] = kEmpty;
        S.vals[i] = 0u;
    }
    if (threadIdx.x == 0) S.qn = 0;
    const float* src = it.tris + first * 9;
    for (int i = threadIdx.x; i < count * 9; i += kChunk) S.tri[i] = src[i];
    __syncthreads();
    if ((int)threadIdx.x < count) {
        float v[3][3];
        staged_vertices(S, threadIdx.x, info, v);
        SurfTri T;
        surface_setup(v, info, T);
        const int nx = T.hi[0] - T.lo[0] + 1;
        const int64_t vol = (int64_t)nx * (T.hi[1] - T.lo[1] + 1) * (T.hi[2] - T.lo[2] + 1);
        bool queued = false;
        if (vol > kBigVox || nx > 32) {
            S.queue[atomicAdd(&S.qn, 1u)] = threadIdx.x;  // at most kChunk entries
            queued = true;
        }
        if (!queued) lane_raster(S, table, T, info, layout);  // nx <= 32 here
    }
    __syncthreads();
    // queued triangles: rows x 32-voxel chunks split over the workgroup
    const unsigned nq = S.qn < (unsigned)kQueue ? S.qn : (unsigned)kQueue;
    for (unsigned q = 0; q < nq; q++) {
        float v[3][3];
        staged_vertices(S, (int)S.queue[q], info, v);
        SurfTri T;
        surface_setup(v, info, T);
        float xy[9], zx[9];
        tri_coeffs(T, xy, zx);
        const int ny = T.hi[1] - T.lo[1] + 1, nz = T.hi[2] - T.lo[2] + 1;
        const int c0 = T.lo[0] >> 5, nch = (T.hi[0] >> 5) - c0 + 1;
        const int64_t items = (int64_t)ny * nz * nch;
        for (int64_t i = threadIdx.x; i < items; i += kChunk) {
            const int ch = (int)(i % nch);
            const int64_t r = i / nch;
            const int y = T.lo[1] + (int)(r % ny), z = T.lo[2] + (int)(r / ny);
            const float py = (float)y * info.unit[1], pz = (float)z * info.unit[2];
            if (!edges_pass(T.yz, py, pz)) continue;
            const int cx = (c0 + ch) << 5;
            const int x0 = cx > T.lo[0] ? cx : T.lo[0], x1 = cx + 31 < T.hi[0] ? cx + 31 : T.hi[0];
            chunk_row(S, table, T, xy, zx, info, layout, x0, x1, cx, y, z, py, pz);
        }
    }
    __syncthreads();  // every hash insert is done
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
    if (M.items)
        surface_chunk_kernel<true><<<units, kChunk, 0, stream>>>(one, M);
    else
        surface_chunk_kernel<false><<<units, kChunk, 0, stream>>>(one, M);
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
