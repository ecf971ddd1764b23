// prep.hip -- the steps either side of the hot path, on the device.
//
//   vox_mesh_bbox        component-wise min/max of the vertices (trimesh2's
//                        need_bbox, main.cpp:179) as a device reduction
//   vox_pack_triangles   indexed mesh -> AoS FP32 triangles (main.cpp:61-80,
//                        meshToGPU_managed) as a gather kernel, with index checks
//   vox_upload_async     host -> device through double-buffered pinned staging
//                        (replaces the reference's managed-memory upload)
//   vox_convert_layout   linear <-> Morton bit tables (voxelize.cuh:20-34 vs
//                        util.h:25-38 / checkVoxel)
//   vox_count_voxels     number of set voxels (popcount reduction)
#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <mutex>
#include <utility>
#include <string>

#include "../../include/cuda_voxelizer_amd/vox_capi.h"
#include "vox_device.h"
#include "vox_host.h"
#include "vox_kernels.h"

namespace voxamd {

// Order-preserving map of a finite float to uint32 (negative values reversed).
__device__ __forceinline__ uint32_t float_key(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_float(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

// Grid-stride min/max keys, reduced over the wave (shuffles) and the block
// (LDS) into the block's six partial keys; the last block to finish (a
// counter) reduces the partials, writes the six floats and returns the
// counter to zero, so one launch does it all.  ws: [0] counter, [64..]
// kBboxBlocks x 6 partial keys.
constexpr unsigned kBboxBlocks = 64;

__global__ void __launch_bounds__(256) bbox_kernel(const float* __restrict__ v, uint64_t n, uint32_t* ws,
                                                   float* __restrict__ out) {
    __shared__ uint32_t part[4][6];
    uint32_t mn[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, mx[3] = {0u, 0u, 0u};
    // eight vertices per step, their 24 loads issued before any is used
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i0 = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i0 < n; i0 += 8 * stride) {
        float f[8][3];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const uint64_t i = i0 + (uint64_t)u * stride;
#pragma unroll
            for (int k = 0; k < 3; k++) f[u][k] = i < n ? v[3 * i + k] : v[3 * i0 + k];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const uint32_t key = float_key(f[u][k]);
                mn[k] = key < mn[k] ? key : mn[k];
                mx[k] = key > mx[k] ? key : mx[k];
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        for (int d = 32; d > 0; d >>= 1) {
            const uint32_t a = __shfl_xor(mn[k], d), b = __shfl_xor(mx[k], d);
            mn[k] = a < mn[k] ? a : mn[k];
            mx[k] = b > mx[k] ? b : mx[k];
        }
    }
    const int wave = threadIdx.x >> 6;
    if (__lane_id() == 0) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            part[wave][k] = mn[k];
            part[wave][3 + k] = mx[k];
        }
    }
    __syncthreads();
    __shared__ bool last;
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        uint32_t r = part[0][k];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) r = k < 3 ? (part[w][k] < r ? part[w][k] : r)
                                                                  : (part[w][k] > r ? part[w][k] : r);
        ws[64 + blockIdx.x * 6 + k] = r;
        __threadfence();
    }
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(ws, 1u) == gridDim.x - 1;
    __syncthreads();
    if (last) {  // uniform over the block
        __shared__ uint32_t allp[kBboxBlocks * 6];
        __threadfence();
        const volatile uint32_t* pk = ws + 64;
        for (unsigned t = threadIdx.x; t < gridDim.x * 6u; t += blockDim.x) allp[t] = pk[t];
        __syncthreads();
        if (threadIdx.x < 6) {
            const int k = threadIdx.x;
            uint32_t r = allp[k];
            for (unsigned b = 1; b < gridDim.x; b++) {
                const uint32_t x = allp[b * 6 + k];
                r = k < 3 ? (x < r ? x : r) : (x > r ? x : r);
            }
            out[k] = key_float(r);
            if (k == 0) ws[0] = 0u;
        }
    }
}

// One workgroup per 256 faces: indices and triangles pass through LDS so the
// global loads and stores are coalesced (a face is 12 bytes in, 36 bytes out);
// each lane gathers its face's three vertices.
__global__ void __launch_bounds__(256) pack_kernel(const float* __restrict__ v, uint64_t nv,
                                                   const int32_t* __restrict__ faces, uint64_t nf,
                                                   float* __restrict__ out, unsigned long long* bad) {
    __shared__ int32_t s_idx[256 * 3];
    __shared__ float s_out[256 * 9];
    unsigned long long nbad = 0;
    for (uint64_t base = blockIdx.x * 256ull; base < nf; base += (uint64_t)gridDim.x * 256ull) {
        const int cnt = (int)(nf - base < 256ull ? nf - base : 256ull);
        const int32_t* fsrc = faces + 3 * base;
        int32_t fi[3];
#pragma unroll
        for (int u = 0; u < 3; u++) {
            const int i = (int)threadIdx.x + 256 * u;
            fi[u] = i < 3 * cnt ? fsrc[i] : 0;
        }
#pragma unroll
        for (int u = 0; u < 3; u++) {
            const int i = (int)threadIdx.x + 256 * u;
            if (i < 3 * cnt) s_idx[i] = fi[u];
        }
        __syncthreads();
        if ((int)threadIdx.x < cnt) {
            float t[9];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const int32_t idx = s_idx[3 * threadIdx.x + c];
                const bool ok = idx >= 0 && (uint64_t)idx < nv;
                if (!ok) nbad++;
                const uint64_t src = ok ? 3 * (uint64_t)idx : 0;
#pragma unroll
                for (int k = 0; k < 3; k++) t[3 * c + k] = ok ? v[src + k] : 0.0f;
            }
#pragma unroll
            for (int k = 0; k < 9; k++) s_out[9 * threadIdx.x + k] = t[k];
        }
        __syncthreads();
        float* dst = out + 9 * base;
#pragma unroll
        for (int u = 0; u < 9; u++) {
            const int i = (int)threadIdx.x + 256 * u;
            if (i < 9 * cnt) dst[i] = s_out[i];
        }
        __syncthreads();  // s_idx / s_out are reused by the next group
    }
    if (bad && nbad) atomicAdd(bad, nbad);
}

// Linear <-> Morton for the grids the block form does not take (G < 32,
// G > 1024, unaligned tables): one thread per destination word gathers its 32 voxels.
__global__ void __launch_bounds__(256) convert_kernel(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                                                      uint32_t G, int to_morton) {
    const uint64_t words = ((uint64_t)G * G * G) >> 5;
    for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < words;
         w += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t out = 0;
        for (int j = 0; j < 32; j++) {
            const uint64_t di = (w << 5) | (uint64_t)j;  // destination voxel index
            uint32_t x, y, z;
            if (to_morton) {  // di is a Morton code
                x = y = z = 0;
                for (int b = 0; b < 16; b++) {
                    x |= (uint32_t)((di >> (3 * b)) & 1u) << b;
                    y |= (uint32_t)((di >> (3 * b + 1)) & 1u) << b;
                    z |= (uint32_t)((di >> (3 * b + 2)) & 1u) << b;
                }
            } else {
                x = (uint32_t)(di % G);
                y = (uint32_t)((di / G) % G);
                z = (uint32_t)(di / ((uint64_t)G * G));
            }
            const uint64_t si = to_morton ? (uint64_t)x + (uint64_t)y * G + (uint64_t)z * G * G : morton3(x, y, z);
            if ((src[si >> 5] >> (31u - (uint32_t)(si & 31u))) & 1u) out |= 1u << (31 - j);
        }
        dst[w] = out;
    }
}

__global__ void __launch_bounds__(256) count_kernel(const uint32_t* __restrict__ t, uint64_t words,
                                                    unsigned long long* out) {
    unsigned long long c = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < words; i += (uint64_t)gridDim.x * blockDim.x)
        c += (unsigned long long)__popc(t[i]);
    for (int d = 32; d > 0; d >>= 1) c += __shfl_xor(c, d);
    if (__lane_id() == 0 && c) atomicAdd(out, c);
}

static unsigned grid_for(uint64_t n) {
    uint64_t b = (n + 255) / 256;
    if (b < 1) b = 1;
    if (b > 4096) b = 4096;
    return (unsigned)b;
}

// pinned staging for vox_upload_async (double buffered, per process)
static std::mutex g_bbox_mutex;
static std::map<std::pair<int, hipStream_t>, void*> g_bbox_ws;
static std::mutex g_up_mutex;
static void* g_up_buf[2] = {nullptr, nullptr};
static hipEvent_t g_up_ev[2] = {nullptr, nullptr};
constexpr size_t kUpChunk = 8u << 20;

}  // namespace voxamd

using namespace voxamd;

extern "C" {

vox_status vox_mesh_bbox(const float* d_vertices, uint64_t n_vertices, float* d_out6, void* stream) {
    if (!d_out6 || (n_vertices && !d_vertices)) return fail(VOX_ERR_INVALID, "NULL argument");
    if (n_vertices == 0) return fail(VOX_ERR_INVALID, "empty vertex array");
    hipStream_t st = (hipStream_t)stream;
    uint32_t* ws = nullptr;
    {
        // per (device, stream): the counter and the partial keys; the counter
        // starts at zero and the kernel's last block returns it there
        int dev = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
        std::lock_guard<std::mutex> lock(g_bbox_mutex);
        void*& w = g_bbox_ws[{dev, st}];
        if (!w) {
            e = hipMalloc(&w, 4u * (64u + 6u * kBboxBlocks));
            if (e == hipSuccess) e = hipMemset(w, 0, 4u * 64u);
            if (e == hipSuccess) e = hipDeviceSynchronize();
            if (e != hipSuccess) {
                if (w) (void)hipFree(w);
                w = nullptr;
                return hip_fail(e, "bbox workspace");
            }
        }
        ws = (uint32_t*)w;
    }
    const unsigned nb = grid_for(n_vertices) < kBboxBlocks ? grid_for(n_vertices) : kBboxBlocks;
    bbox_kernel<<<nb, 256, 0, st>>>(d_vertices, n_vertices, ws, d_out6);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? VOX_OK : hip_fail(e, "bbox kernels");
}

vox_status vox_pack_triangles(const float* d_vertices, uint64_t n_vertices, const int32_t* d_faces,
                              uint64_t n_faces, float* d_tris, uint64_t* d_bad, void* stream) {
    if (n_faces && (!d_vertices || !d_faces || !d_tris)) return fail(VOX_ERR_INVALID, "NULL argument");
    if (n_faces == 0) return VOX_OK;
    pack_kernel<<<grid_for(n_faces), 256, 0, (hipStream_t)stream>>>(d_vertices, n_vertices, d_faces, n_faces, d_tris,
                                                                    (unsigned long long*)d_bad);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? VOX_OK : hip_fail(e, "pack kernel");
}

vox_status vox_upload_async(const void* host, size_t bytes, void* d_dst, void* stream) {
    if (bytes && (!host || !d_dst)) return fail(VOX_ERR_INVALID, "NULL argument");
    hipStream_t st = (hipStream_t)stream;
    std::lock_guard<std::mutex> lock(g_up_mutex);
    hipError_t e = hipSuccess;
    for (int i = 0; i < 2 && e == hipSuccess; i++) {
        if (!g_up_buf[i]) e = hipHostMalloc(&g_up_buf[i], kUpChunk, hipHostMallocDefault);
        if (e == hipSuccess && !g_up_ev[i]) e = hipEventCreateWithFlags(&g_up_ev[i], hipEventDisableTiming);
    }
    if (e != hipSuccess) return hip_fail(e, "pinned staging");
    size_t off = 0;
    int k = 0;
    while (off < bytes) {
        const size_t n = bytes - off < kUpChunk ? bytes - off : kUpChunk;
        e = hipEventSynchronize(g_up_ev[k]);  // the previous copy out of this buffer is done
        if (e != hipSuccess) return hip_fail(e, "hipEventSynchronize");
        std::memcpy(g_up_buf[k], (const char*)host + off, n);
        e = hipMemcpyAsync((char*)d_dst + off, g_up_buf[k], n, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipEventRecord(g_up_ev[k], st);
        if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync");
        off += n;
        k ^= 1;
    }
    return VOX_OK;
}

vox_status vox_convert_layout(const uint32_t* d_src, uint32_t* d_dst, uint32_t grid, int to_morton, void* stream) {
    if (!d_src || !d_dst || d_src == d_dst) return fail(VOX_ERR_INVALID, "need distinct source and destination");
    if (grid < 4 || (grid & (grid - 1)) || grid > kMaxGrid)
        return fail(VOX_ERR_UNSUPPORTED, "layout conversion needs a power-of-two grid >= 4");
    const uint64_t words = ((uint64_t)grid * grid * grid) >> 5;
    hipError_t e;
    if (layout_convert_fast(grid) && ((((uintptr_t)d_src) | ((uintptr_t)d_dst)) & 15u) == 0) {
        // blocks of 32 x 8 rows through LDS, 8 Morton words per thread (marker.hip)
        e = launch_layout_convert(d_src, d_dst, grid, to_morton != 0, (hipStream_t)stream);
    } else {
        convert_kernel<<<grid_for(words), 256, 0, (hipStream_t)stream>>>(d_src, d_dst, grid, to_morton ? 1 : 0);
        e = hipGetLastError();
    }
    return e == hipSuccess ? VOX_OK : hip_fail(e, "convert kernel");
}

vox_status vox_count_voxels(const uint32_t* d_table, uint64_t words, uint64_t* d_count, void* stream) {
    if (!d_count || (words && !d_table)) return fail(VOX_ERR_INVALID, "NULL argument");
    if (words == 0) return VOX_OK;
    count_kernel<<<grid_for(words), 256, 0, (hipStream_t)stream>>>(d_table, words, (unsigned long long*)d_count);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? VOX_OK : hip_fail(e, "count kernel");
}

int vox_check_voxel(const uint32_t* table, uint32_t grid, uint32_t x, uint32_t y, uint32_t z, int layout) {
    if (!table || x >= grid || y >= grid || z >= grid) return 0;
    uint64_t idx;
    if (layout == VOX_LAYOUT_MORTON) {
        idx = 0;
        for (int b = 0; b < 21; b++)
            idx |= ((uint64_t)((x >> b) & 1u) << (3 * b)) | ((uint64_t)((y >> b) & 1u) << (3 * b + 1)) |
                   ((uint64_t)((z >> b) & 1u) << (3 * b + 2));
    } else {
        idx = (uint64_t)x + (uint64_t)y * grid + (uint64_t)z * grid * grid;
    }
    return (int)((table[idx >> 5] >> (31u - (uint32_t)(idx & 31u))) & 1u);
}

}  // extern "C"
