// region.hip -- the region algorithm's build pass, plan and workspace.
//
// Phase 1 (chunk.hip surface_chunk_kernel<..., REG>, marker.hip
// solid_marker_kernel<..., REG>) stages (key, mask) pairs per output region;
// the build pass here assembles each region in LDS from its pairs (and from
// the scratch words of a dirty region), takes the row suffix XOR for solid
// tables, and writes every word of the region once.  Regions that received
// nothing are streamed as zeros without touching LDS.
#include "bricks.h"
#include "region.h"
#include "vox_kernels.h"

#include <map>
#include <mutex>
#include <utility>

namespace voxamd {

constexpr uint32_t kRegionDcap = 512;   // descriptors per region (more: the pairs fall back)
constexpr uint32_t kRegionSlice = 2048; // staged pairs per phase-1 workgroup: its whole LDS hash
constexpr int kBuildThreads = 256;

__device__ __forceinline__ uint32_t region_lds_addr(int mode, int lrw, uint32_t local) {
    if (mode == REGION_BRICKS) return (local >> lrw) * ((1u << lrw) + 1u) + (local & ((1u << lrw) - 1u));
    return local;
}

// Table word of local word i of a ROWS region whose first row is (y0, z0).
__device__ __forceinline__ uint64_t rows_table_word(const RegionPlan& P, uint32_t y0, uint32_t z0, uint32_t i) {
    const uint32_t row = i >> P.lrw, xw = i & ((1u << P.lrw) - 1u);
    const uint32_t zl = row >> P.lby, yl = row & ((1u << P.lby) - 1u);
    return ((((uint64_t)(z0 + zl) << P.lgG) + y0 + yl) << P.lrw) + xw;
}

// One workgroup per region.
template <int MODE>
__global__ void __launch_bounds__(kBuildThreads)
    region_build_kernel(RegionPlan P, RegionWs W, uint32_t* __restrict__ table, int accumulate) {
    extern __shared__ uint32_t buf[];
    __shared__ uint32_t s_off[kRegionDcap];
    __shared__ uint32_t s_pre[kRegionDcap + 1];
    __shared__ uint32_t s_wsum[kBuildThreads / 64];
    __shared__ uint32_t s_n, s_dirty;
    const uint32_t r = blockIdx.x;
    const uint32_t words = 1u << P.lg;
    const uint32_t rw = 1u << P.lrw;
    if (threadIdx.x == 0) {
        const uint32_t c = W.rdc[r];
        const uint32_t d = W.dirty[r];
        s_n = c < P.dcap ? c : P.dcap;
        s_dirty = d;
        if (c) W.rdc[r] = 0u;    // the lists start empty at the next call
        if (d) W.dirty[r] = 0u;
    }
    __syncthreads();
    const uint32_t n = s_n;
    const bool dirty = s_dirty != 0u;
    // region placement: RANGE -> words [r << lg, +words); ROWS / BRICKS -> rows from (y0, z0)
    const uint32_t yblocks_lg = (uint32_t)(P.lgG - P.lby);
    const uint32_t y0 = (r & ((1u << yblocks_lg) - 1u)) << P.lby;
    const uint32_t z0 = (r >> yblocks_lg) << P.lbz;
    const bool aligned = (((uintptr_t)table) & 15u) == 0;
    if (MODE != REGION_BRICKS && n == 0 && !dirty && !accumulate) {
        // nothing staged: zeros, straight to memory
        if (MODE == REGION_RANGE) {
            const uint64_t w0 = (uint64_t)r << P.lg;
            if (aligned) {
                uint4* t4 = (uint4*)(table + w0);
                for (uint32_t i = threadIdx.x; i < (words >> 2); i += kBuildThreads) t4[i] = make_uint4(0u, 0u, 0u, 0u);
            } else {
                for (uint32_t i = threadIdx.x; i < words; i += kBuildThreads) table[w0 + i] = 0u;
            }
        } else {
            const uint32_t seg = 1u << (P.lby + P.lrw);  // words of one z row block (contiguous)
            for (uint32_t zl = 0; zl < (1u << P.lbz); zl++) {
                const uint64_t w0 = ((((uint64_t)(z0 + zl) << P.lgG) + y0) << P.lrw);
                if (aligned && (seg & 3u) == 0) {
                    uint4* t4 = (uint4*)(table + w0);
                    for (uint32_t i = threadIdx.x; i < (seg >> 2); i += kBuildThreads)
                        t4[i] = make_uint4(0u, 0u, 0u, 0u);
                } else {
                    for (uint32_t i = threadIdx.x; i < seg; i += kBuildThreads) table[w0 + i] = 0u;
                }
            }
        }
        return;
    }
    // 1. the region's scratch words (dirty) or zeros
    const uint64_t key0 = (uint64_t)r << P.lg;
    for (uint32_t i = threadIdx.x; i < words; i += kBuildThreads) {
        uint32_t v = 0u;
        if (dirty) {
            v = W.scratch[key0 + i];
            if (v) W.scratch[key0 + i] = 0u;
        }
        buf[region_lds_addr(MODE, P.lrw, i)] = v;
    }
    // 2. descriptor offsets and the exclusive prefix of their counts
    {
        const uint32_t d0 = 2u * threadIdx.x, d1 = d0 + 1u;
        uint32_t c0 = 0u, c1 = 0u;
        if (d0 < n) {
            const uint2 e = W.desc[(uint64_t)r * P.dcap + d0];
            s_off[d0] = e.x;
            c0 = e.y;
        }
        if (d1 < n) {
            const uint2 e = W.desc[(uint64_t)r * P.dcap + d1];
            s_off[d1] = e.x;
            c1 = e.y;
        }
        const int lane = __lane_id(), wave = threadIdx.x >> 6;
        uint32_t incl = c0 + c1;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t o = __shfl_up(incl, d);
            if (lane >= d) incl += o;
        }
        if (lane == 63) s_wsum[wave] = incl;
        __syncthreads();
        uint32_t before = 0u, total = 0u;
#pragma unroll
        for (int w = 0; w < kBuildThreads / 64; w++) {
            if (w < wave) before += s_wsum[w];
            total += s_wsum[w];
        }
        const uint32_t excl = before + incl - c0 - c1;
        if (d0 < n) s_pre[d0] = excl;
        if (d1 < n) s_pre[d1] = excl + c0;
        if (threadIdx.x == 0) s_pre[n] = total;
    }
    __syncthreads();  // buf initialised, s_off / s_pre complete
    // 3. the pairs: item j of the concatenated lists is pair j - s_pre[d] of list d
    const uint32_t total = s_pre[n];
    const uint32_t lmask = words - 1u;
    uint32_t d = 0;
    for (uint32_t j0 = threadIdx.x; j0 < total; j0 += 4u * kBuildThreads) {
        uint2 p[4];
        bool have[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint32_t j = j0 + (uint32_t)u * kBuildThreads;
            have[u] = j < total;
            if (have[u]) {
                while (d + 1u < n && s_pre[d + 1u] <= j) d++;
                p[u] = W.stage[s_off[d] + (j - s_pre[d])];
            }
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (!have[u]) continue;
            uint32_t* a = buf + region_lds_addr(MODE, P.lrw, p[u].x & lmask);
            if (P.solid)
                atomicXor(a, p[u].y);
            else
                atomicOr(a, p[u].y);
        }
    }
    __syncthreads();
    // 4. solid: row suffix; write-out
    if (MODE == REGION_BRICKS) {
        const uint32_t ts = rw + 1u, nrows = words >> P.lrw;
        if (P.solid) {
            tile_suffix(buf, ts, rw, nrows, [&](uint32_t row, uint32_t j, uint32_t fin) { buf[row * ts + j] = fin; });
            __syncthreads();
        }
        tile_to_bricks(buf, ts, table, P.G, rw, 1u << P.lbz, y0, z0, accumulate);
        return;
    }
    if (MODE == REGION_ROWS && P.solid) {  // suffix in place, then the 16-byte write-out below
        tile_suffix(buf, rw, rw, words >> P.lrw, [&](uint32_t row, uint32_t j, uint32_t fin) { buf[row * rw + j] = fin; });
        __syncthreads();
    }
    // RANGE is one contiguous range; ROWS is 2^lbz contiguous row blocks
    const uint32_t seg_lg = MODE == REGION_RANGE ? (uint32_t)P.lg : (uint32_t)(P.lby + P.lrw);
    const uint32_t seg = 1u << seg_lg;
    if (aligned && (seg & 3u) == 0) {
        for (uint32_t i4 = threadIdx.x; i4 < (words >> 2); i4 += kBuildThreads) {
            const uint32_t i = i4 << 2;
            const uint64_t w = MODE == REGION_RANGE ? key0 + i : rows_table_word(P, y0, z0, i);
            uint4 v = make_uint4(buf[i], buf[i + 1], buf[i + 2], buf[i + 3]);
            uint4* t4 = (uint4*)(table + w);
            if (accumulate) {  // OR (surface) / XOR (solid) into the caller's table
                const uint4 o = *t4;
                v = P.solid ? make_uint4(v.x ^ o.x, v.y ^ o.y, v.z ^ o.z, v.w ^ o.w)
                            : make_uint4(v.x | o.x, v.y | o.y, v.z | o.z, v.w | o.w);
            }
            *t4 = v;
        }
    } else {
        for (uint32_t i = threadIdx.x; i < words; i += kBuildThreads) {
            const uint64_t w = MODE == REGION_RANGE ? key0 + i : rows_table_word(P, y0, z0, i);
            table[w] = accumulate ? (P.solid ? (table[w] ^ buf[i]) : (table[w] | buf[i])) : buf[i];
        }
    }
}

// Region shapes.  Power-of-two G in 32..2048 (the descriptor lists grow
// with the region count: 128 MiB at 2048^3).  lg_max: log2 words of the
// largest region (LDS per build workgroup).
bool make_region_plan(uint32_t G, int layout, bool solid, RegionPlan& P, const RegionKnobs& K) {
    if (G < 32u || G > 2048u || (G & (G - 1u))) return false;
    const int lg_max = K.lg_words >= 10 && K.lg_words <= 13 ? K.lg_words : 13;
    int lgG = 0;
    while ((1u << lgG) < G) lgG++;
    const int lwords = 3 * lgG - 5;
    P = RegionPlan{};
    P.G = G;
    P.lgG = lgG;
    P.lrw = lgG - 5;
    P.solid = solid ? 1 : 0;
    if (layout == VOX_LAYOUT_MORTON && solid) {
        // the brick writer's row blocks: 32 y x rz z, all x
        P.mode = REGION_BRICKS;
        P.lby = 5;
        P.lbz = G <= 1024u ? 3 : 1;
        P.lg = P.lrw + P.lby + P.lbz;
    } else if (layout == VOX_LAYOUT_MORTON) {
        P.mode = REGION_RANGE;
        P.lg = lwords < lg_max ? lwords : lg_max;
        P.lby = 0;
        P.lbz = 0;
    } else {
        P.mode = REGION_ROWS;
        P.lg = lwords < lg_max ? lwords : lg_max;
        const int rows_lg = P.lg - P.lrw;  // rows per region
        if (rows_lg < 0) return false;
        P.lby = rows_lg < 6 ? rows_lg : 6;
        if (P.lby > lgG) P.lby = lgG;
        P.lbz = rows_lg - P.lby;
        if (P.lbz > lgG) return false;
    }
    if (P.lg > 14 || P.lg > lwords) return false;
    P.nreg = 1u << (lwords - P.lg);
    P.dcap = K.lg_desc >= 0 && K.lg_desc <= 9 ? 1u << K.lg_desc : kRegionDcap;
    P.slice = K.lg_stage >= 0 && K.lg_stage < 11 ? 1u << K.lg_stage : kRegionSlice;
    return true;
}

// ---- per (device, stream) workspace -------------------------------------------
// [rdc][dirty][desc][stage][scratch]; rdc, dirty and the scratch table are
// zero-filled when allocated, and every call leaves them zero.
struct RegionBuffer {
    void* ptr = nullptr;
    uint64_t nreg_cap = 0, desc_cap = 0, scap = 0, scratch_words = 0, bytes = 0;
};
static std::mutex g_region_mutex;
static std::map<std::pair<int, hipStream_t>, RegionBuffer> g_region;

static uint64_t align256(uint64_t b) { return (b + 255u) & ~255ull; }

hipError_t region_workspace(const RegionPlan& P, uint64_t workgroups, hipStream_t stream, RegionWs& W) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    const uint64_t need_nreg = P.nreg, need_desc = (uint64_t)P.nreg * P.dcap;
    const uint64_t need_stage = (workgroups ? workgroups : 1) * (uint64_t)P.slice;
    const uint64_t need_scratch = (uint64_t)P.nreg << P.lg;
    std::lock_guard<std::mutex> lock(g_region_mutex);
    RegionBuffer& B = g_region[std::make_pair(dev, stream)];
    if (!B.ptr || B.nreg_cap < need_nreg || B.desc_cap < need_desc || B.scap < need_stage ||
        B.scratch_words < need_scratch) {
        RegionBuffer nb;
        nb.nreg_cap = B.nreg_cap > need_nreg ? B.nreg_cap : need_nreg;
        nb.desc_cap = B.desc_cap > need_desc ? B.desc_cap : need_desc;
        nb.scap = B.scap > need_stage ? B.scap : need_stage;
        nb.scratch_words = B.scratch_words > need_scratch ? B.scratch_words : need_scratch;
        const uint64_t zero_bytes = align256(4 * nb.nreg_cap) * 2;
        nb.bytes = zero_bytes + align256(8 * nb.desc_cap) + align256(8 * nb.scap) + align256(4 * nb.scratch_words);
        if (B.ptr) {
            e = hipStreamSynchronize(stream);  // a previous call on this stream may still use it
            if (e != hipSuccess) return e;
            (void)hipFree(B.ptr);
            B = RegionBuffer{};
        }
        e = hipMalloc(&nb.ptr, nb.bytes);
        if (e != hipSuccess) return e;
        e = hipMemset(nb.ptr, 0, zero_bytes);
        if (e == hipSuccess) {
            const uint64_t scratch_off = zero_bytes + align256(8 * nb.desc_cap) + align256(8 * nb.scap);
            e = hipMemset((char*)nb.ptr + scratch_off, 0, 4 * nb.scratch_words);
        }
        if (e == hipSuccess) e = hipDeviceSynchronize();  // zero before any stream's first use
        if (e != hipSuccess) {
            (void)hipFree(nb.ptr);
            return e;
        }
        B = nb;
    }
    char* p = (char*)B.ptr;
    W.rdc = (unsigned*)p;
    p += align256(4 * B.nreg_cap);
    W.dirty = (unsigned*)p;
    p += align256(4 * B.nreg_cap);
    W.desc = (uint2*)p;
    p += align256(8 * B.desc_cap);
    W.stage = (uint2*)p;
    p += align256(8 * B.scap);
    W.scratch = (uint32_t*)p;
    return hipSuccess;
}

hipError_t launch_region_build(const RegionPlan& P, const RegionWs& W, uint32_t* table, bool accumulate,
                               hipStream_t stream) {
    const uint32_t words = 1u << P.lg;
    const size_t lds = 4ull * (P.mode == REGION_BRICKS ? (words >> P.lrw) * ((1u << P.lrw) + 1u) : words);
    if (P.mode == REGION_RANGE)
        region_build_kernel<REGION_RANGE><<<P.nreg, kBuildThreads, lds, stream>>>(P, W, table, accumulate ? 1 : 0);
    else if (P.mode == REGION_ROWS)
        region_build_kernel<REGION_ROWS><<<P.nreg, kBuildThreads, lds, stream>>>(P, W, table, accumulate ? 1 : 0);
    else
        region_build_kernel<REGION_BRICKS><<<P.nreg, kBuildThreads, lds, stream>>>(P, W, table, accumulate ? 1 : 0);
    return hipGetLastError();
}

bool region_supported(uint32_t G, int layout, bool solid) {
    RegionPlan P;
    return make_region_plan(G, layout, solid, P, RegionKnobs{});
}

hipError_t launch_surface_region(const vox_info& info, const float* tris, uint32_t* table, int layout,
                                 bool accumulate, hipStream_t stream, uint32_t big, int waves, const RegionKnobs& K) {
    RegionPlan P;
    if (!make_region_plan(info.gridsize[0], layout, false, P, K)) return hipErrorInvalidValue;
    RegionWs W;
    hipError_t e = region_workspace(P, chunk_units(info.n_triangles, 1, 1), stream, W);
    if (e != hipSuccess) return e;
    e = launch_surface_region_chunks(info, tris, layout, RegionArgs{P, W}, stream, big, waves);
    if (e != hipSuccess) return e;
    return launch_region_build(P, W, table, accumulate, stream);
}

}  // namespace voxamd
