// tile.hip -- binned, tile-owned voxelization kernels for gfx950.
//
// Pipeline per call (all on one stream, no host synchronisation):
//   1. bin_count   : per triangle, the output tiles its voxel (surface) or
//                    column (solid) range touches; wave-aggregated atomic
//                    counts per tile.  Triangles touching more than
//                    P.max_tiles_per_tri tiles go to the wide list instead.
//   2. bin_scan    : exclusive scan of the counts (one workgroup).
//   3. bin_scatter : triangle ids into per-tile lists (wave-aggregated slots).
//   4. raster_*    : one workgroup per tile builds the tile in LDS (ds_or /
//                    ds_xor), then writes every word of the tile once with
//                    plain coalesced stores -- this is also the zero-fill.
//                    Lanes take one triangle each; triangles whose clipped
//                    voxel box is large are queued and split across the
//                    whole workgroup.
//   5. wide_*      : triangles spanning many tiles (and, for the solid
//                    linear layout, runs that spill past the end of a row)
//                    are applied afterwards with global atomics, split
//                    across a workgroup per triangle.
// The workspace is self-cleaning: the raster pass zeroes the counts it used
// and the last wide block zeroes the counters, so the next call on the same
// workspace needs no memset (the pipeline can be captured in a HIP graph).
//
// The per-voxel and per-column arithmetic is vox_device.h's, i.e. the
// reference CPU path's (cpu_voxelizer.cpp:56-171, :255-310) bit for bit.
#include "jobs.h"
#include "raster.h"
#include "tile.h"
#include "vox_device.h"
#include "vox_kernels.h"

#pragma clang fp contract(off)

namespace voxamd {

constexpr int kBlock = 256;
constexpr int kWideBlocks = 128;  // persistent grid of the wide pass

// misc counter slots (uint32 words in the workspace)
// MISC_WIDE counts wide triangles during bin_count; bin_scan copies it to
// MISC_WIDE_N (read by the wide pass) and resets MISC_WIDE and the spill
// counters, which the raster pass then fills -- so every counter is zero
// again before the next call's bin_count without any fence or extra launch.
enum : int { MISC_WIDE = 0, MISC_SPILL = 1, MISC_SPILL_OVF = 2, MISC_WIDE_N = 3, MISC_WORDS = 8 };

struct Range {  // inclusive tile-index ranges
    int a0, a1, b0, b1, c0, c1;
    __device__ int count() const { return (a1 - a0 + 1) * (b1 - b0 + 1) * (c1 - c0 + 1); }
};

__device__ __forceinline__ int tile_id(const TilePlan& P, int a, int b, int c) {
    if (P.kind == TILE_MORTON_CUBE) return (int)morton3((uint32_t)a, (uint32_t)b, (uint32_t)c);
    return b + c * P.nty;  // row tiles: a (x) is always 0
}

__device__ __forceinline__ int range_tile(const TilePlan& P, const Range& R, int k) {
    const int na = R.a1 - R.a0 + 1, nb = R.b1 - R.b0 + 1;
    const int a = R.a0 + k % na;
    const int b = R.b0 + (k / na) % nb;
    const int c = R.c0 + k / (na * nb);
    return tile_id(P, a, b, c);
}

// Tile range of a surface triangle from its clamped voxel box.
__device__ __forceinline__ Range surface_range(const TilePlan& P, const SurfTri& T) {
    Range R;
    R.a0 = T.lo[0] / P.tx; R.a1 = T.hi[0] / P.tx;
    R.b0 = T.lo[1] / P.ty; R.b1 = T.hi[1] / P.ty;
    R.c0 = T.lo[2] / P.tz; R.c1 = T.hi[2] / P.tz;
    return R;
}

// Column range of a solid triangle, clamped to the grid; false when empty.
// Per-triangle out-of-grid events are counted here (once per triangle).
__device__ __forceinline__ bool solid_columns(const SolidTri& S, int G, int& ylo, int& yhi, int& zlo, int& zhi,
                                              unsigned long long* offrow_events) {
    ylo = (int)S.ymin;
    zlo = (int)S.zmin;
    if (offrow_events) {
        if (ylo < 0) (*offrow_events)++;
        if (zlo < 0) (*offrow_events)++;
        if (S.ymax > (float)(G - 1)) (*offrow_events)++;
        if (S.zmax > (float)(G - 1)) (*offrow_events)++;
    }
    if (ylo < 0) ylo = 0;
    if (zlo < 0) zlo = 0;
    // (float)y <= ymax for integer-valued ymax <=> y <= ymax
    yhi = S.ymax >= (float)(G - 1) ? G - 1 : (S.ymax >= 0.0f ? (int)S.ymax : -1);
    zhi = S.zmax >= (float)(G - 1) ? G - 1 : (S.zmax >= 0.0f ? (int)S.zmax : -1);
    return ylo <= yhi && zlo <= zhi;
}

// Wave-aggregated atomicAdd(ctr + key, 1) over runs of consecutive active
// lanes with equal keys (triangles are spatially coherent, so neighbouring
// lanes usually share a tile): the first lane of each run adds the run
// length, all runs in one instruction, and each lane gets base + its rank.
// Must be called by every lane of the wave (uniform control flow).
template <bool RESULT>
__device__ __forceinline__ unsigned run_add(unsigned* ctr, int key, bool active) {
    const int lane = __lane_id();
    const unsigned long long act = __ballot(active);
    const int prev_key = __shfl_up(key, 1);
    const bool prev_act = lane > 0 && ((act >> (lane - 1)) & 1ull);
    const bool head = active && !(prev_act && prev_key == key);
    const unsigned long long heads = __ballot(head);
    const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    const unsigned long long hb = heads & upto;
    const int my_head = hb ? 63 - __clzll((long long)hb) : lane;
    unsigned base = 0;
    if (head) {
        const unsigned long long breaks = (heads | ~act) & ~upto;  // first break above me ends my run
        const int end = breaks ? __ffsll((long long)breaks) - 1 : 64;
        const unsigned len = (unsigned)(end - lane);
        if (RESULT)
            base = atomicAdd(ctr + key, len);
        else
            atomicAdd(ctr + key, len);
    }
    if (!RESULT) return 0u;
    base = __shfl(base, my_head);
    return base + (unsigned)(lane - my_head);
}

// Stage kBlock consecutive triangles (9 floats each) through LDS with
// coalesced loads; returns this lane's translated vertices.
__device__ __forceinline__ void staged_tri(const float* __restrict__ tris, uint64_t base, uint64_t n, float* s_tri,
                                          const vox_info& info, float v[3][3]) {
    const uint64_t cnt = (n - base) < (uint64_t)kBlock ? (n - base) : (uint64_t)kBlock;
    const float* src = tris + base * 9;
    __syncthreads();
    stage_floats<kBlock>(s_tri, src, (int)cnt * 9);
    __syncthreads();
    if (threadIdx.x < cnt) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            v[k][0] = s_tri[threadIdx.x * 9 + 3 * k + 0] - info.bbox_min[0];
            v[k][1] = s_tri[threadIdx.x * 9 + 3 * k + 1] - info.bbox_min[1];
            v[k][2] = s_tri[threadIdx.x * 9 + 3 * k + 2] - info.bbox_min[2];
        }
    }
}

// ---------------------------------------------------------------------------
// 1 + 3: binning
// ---------------------------------------------------------------------------
// Work unit: kBlock consecutive triangles of one job.  Tile ids are global
// (job tile_base + local tile), list and wide entries are global triangle ids.
template <bool SOLID, bool SCATTER>
__global__ void __launch_bounds__(kBlock) bin_kernel(ItemDev one, ItemMap M, unsigned* __restrict__ counts,
                                                      unsigned* __restrict__ list, unsigned* __restrict__ misc,
                                                      unsigned* __restrict__ wide_ids) {
    __shared__ float s_tri[kBlock * 9];
    __shared__ BlockJob J;
    resolve_job(one, M, blockIdx.x, J);
    if (J.local == ~0u) return;
    const vox_info info = J.it.info;
    const TilePlan P = J.it.P;
    const uint64_t n = info.n_triangles;
    const uint64_t base = (uint64_t)J.local * kBlock;
    if (base >= n) return;
    unsigned long long events = 0;
    const uint64_t t = base + threadIdx.x;
    const bool active = t < n;
    float v[3][3];
    staged_tri(J.it.tris, base, n, s_tri, info, v);
    Range R{0, -1, 0, -1, 0, -1};
    if (active) {
        if (SOLID) {
            SolidTri S;
            solid_setup(v, info, S);
            int ylo, yhi, zlo, zhi;
            if (S.valid && solid_columns(S, P.G, ylo, yhi, zlo, zhi, SCATTER ? nullptr : &events)) {
                R.a0 = 0; R.a1 = 0;
                R.b0 = ylo / P.ty; R.b1 = yhi / P.ty;
                R.c0 = zlo / P.tz; R.c1 = zhi / P.tz;
            }
        } else {
            SurfTri T;
            surface_setup(v, info, T);
            R = surface_range(P, T);
        }
    }
    const unsigned gt = (unsigned)(J.it.tri_base + t);
    const int nt = (R.a1 < R.a0) ? 0 : R.count();
    const bool wide = nt > P.max_tiles_per_tri;
    if (!SCATTER && wide) wide_ids[atomicAdd(misc + MISC_WIDE, 1u)] = gt;
    unsigned* job_counts = counts + J.it.tile_base;
    for (int k = 0; k < P.max_tiles_per_tri; k++) {
        const bool act = active && !wide && k < nt;
        if (!__ballot(act)) break;
        const int key = act ? range_tile(P, R, k) : 0;
        const unsigned slot = run_add<SCATTER>(job_counts, key, act);
        if (SCATTER && act) list[slot] = gt;
    }
    if (!SCATTER && SOLID && J.it.offrow && events) atomicAdd(J.it.offrow, events);
}

// ---------------------------------------------------------------------------
// 2: exclusive scan of counts -> offsets[0..ntiles], counts <- offsets (cursor)
// ---------------------------------------------------------------------------
// Exclusive scan of counts[0 .. nt) into offsets (and the scatter cursor).
// The counts region is zero-padded to a multiple of 4096: per pass each of the
// 1024 threads loads up to 16 uint4 of consecutive counts at once (65536
// counts per pass, one memory latency), passes chain through a carry.
__global__ void __launch_bounds__(1024) bin_scan_kernel(unsigned nt, unsigned* __restrict__ counts,
                                                        unsigned* __restrict__ offsets, unsigned* __restrict__ misc) {
    __shared__ unsigned wave_sums[16];
    __shared__ unsigned s_carry;
    const int tid = threadIdx.x;
    if (tid == 0) {
        misc[MISC_WIDE_N] = misc[MISC_WIDE];
        misc[MISC_WIDE] = 0u;
        misc[MISC_SPILL] = 0u;
        misc[MISC_SPILL_OVF] = 0u;
        s_carry = 0u;
    }
    __syncthreads();
    for (unsigned pass0 = 0; pass0 < (nt > 0 ? nt : 1u); pass0 += 65536u) {
        const unsigned rem = nt - pass0 < 65536u ? nt - pass0 : 65536u;
        const int per4 = (int)((rem + 4095u) / 4096u);  // uint4 per thread, <= 16
        const unsigned beg = pass0 + (unsigned)tid * (unsigned)per4 * 4u;
        uint4 v[16];
        unsigned local = 0;
#pragma unroll
        for (int q = 0; q < 16; q++) {
            if (q < per4) {
                v[q] = *(const uint4*)(counts + beg + 4 * q);
                local += v[q].x + v[q].y + v[q].z + v[q].w;
            }
        }
        const int lane = tid & 63, wave = tid >> 6;
        unsigned x = local;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const unsigned y = __shfl_up(x, d);
            if (lane >= d) x += y;
        }
        if (lane == 63) wave_sums[wave] = x;
        __syncthreads();
        if (tid < 64) {
            const unsigned w = tid < 16 ? wave_sums[tid] : 0u;
            unsigned e = w;
#pragma unroll
            for (int d = 1; d < 16; d <<= 1) {
                const unsigned y = __shfl_up(e, d);
                if (tid >= d) e += y;
            }
            if (tid < 16) wave_sums[tid] = e - w;
        }
        __syncthreads();
        unsigned run = s_carry + x - local + wave_sums[wave];
#pragma unroll
        for (int q = 0; q < 16; q++) {
            if (q < per4) {
                const unsigned c[4] = {v[q].x, v[q].y, v[q].z, v[q].w};
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const unsigned i = beg + 4u * q + j;
                    if (i < nt) {
                        offsets[i] = run;
                        counts[i] = run;  // scatter cursor
                    }
                    run += c[j];
                }
            }
        }
        __syncthreads();
        if (tid == 1023) s_carry = run;
        __syncthreads();
    }
    if (tid == 0) offsets[nt] = s_carry;
}

// ---------------------------------------------------------------------------
// 4a: surface raster, one workgroup per tile
// ---------------------------------------------------------------------------
struct TileBox {
    int x0, y0, z0, x1, y1, z1;  // inclusive voxel bounds of the tile
};

__device__ __forceinline__ TileBox tile_box(const TilePlan& P, int tile) {
    TileBox B;
    if (P.kind == TILE_MORTON_CUBE) {
        // inverse of morton3 on the tile id (ntiles <= 2^20: 7 bits per axis)
        const uint32_t id = (uint32_t)tile;
        uint32_t a = 0, b = 0, c = 0;
        for (int i = 0; i < 10; i++) {
            a |= ((id >> (3 * i)) & 1u) << i;
            b |= ((id >> (3 * i + 1)) & 1u) << i;
            c |= ((id >> (3 * i + 2)) & 1u) << i;
        }
        B.x0 = a * P.tx; B.y0 = b * P.ty; B.z0 = c * P.tz;
    } else {
        B.x0 = 0;
        B.y0 = (tile % P.nty) * P.ty;
        B.z0 = (tile / P.nty) * P.tz;
    }
    B.x1 = B.x0 + P.tx - 1;
    B.y1 = B.y0 + P.ty - 1;
    B.z1 = B.z0 + P.tz - 1;
    return B;
}

// bits 0..4 of v spread to bits 0, 3, 6, 9, 12 (local Morton inside a cube tile, edge <= 32)
__device__ __forceinline__ uint32_t spread5(uint32_t v) {
    return (v & 1u) | ((v & 2u) << 2) | ((v & 4u) << 4) | ((v & 8u) << 6) | ((v & 16u) << 8);
}

// LDS word + bit of voxel (x,y,z) inside tile B
__device__ __forceinline__ void lds_pos(const TilePlan& P, const TileBox& B, int x, int y, int z, int& word,
                                        uint32_t& bit) {
    if (P.kind == TILE_MORTON_CUBE) {
        const uint32_t local = spread5((uint32_t)(x - B.x0)) | (spread5((uint32_t)(y - B.y0)) << 1) |
                               (spread5((uint32_t)(z - B.z0)) << 2);
        word = (int)(local >> 5);
        bit = 1u << (31u - (local & 31u));
    } else {
        const int row = (y - B.y0) + (z - B.z0) * P.ty;
        word = row * P.row_words + (x >> 5);
        bit = 1u << (31u - (uint32_t)(x & 31));
    }
}

// Test voxels x in [xa, xb] of row (y, z) and OR the hits into the LDS tile.
__device__ __forceinline__ void surface_row(const TilePlan& P, const TileBox& B, const SurfTri& T,
                                            const vox_info& info, uint32_t* lds, int xa, int xb, int y, int z,
                                            float py, float pz) {
    int cur = -1;
    uint32_t mask = 0;
    for (int x = xa; x <= xb; x++) {
        const float px = (float)x * info.unit[0];
        if (!(plane_pass(T, px, py, pz) && edges_pass(T.xy, px, py) && edges_pass(T.zx, pz, px))) continue;
        int w;
        uint32_t b;
        lds_pos(P, B, x, y, z, w, b);
        if (w != cur) {
            if (mask) atomicOr(lds + cur, mask);
            cur = w;
            mask = 0;
        }
        mask |= b;
    }
    if (mask) atomicOr(lds + cur, mask);
}

__device__ __forceinline__ void clip_box(const SurfTri& T, const TileBox& B, int& xa, int& xb, int& ya, int& yb,
                                         int& za, int& zb) {
    xa = max(T.lo[0], B.x0); xb = min(T.hi[0], B.x1);
    ya = max(T.lo[1], B.y0); yb = min(T.hi[1], B.y1);
    za = max(T.lo[2], B.z0); zb = min(T.hi[2], B.z1);
}

// Flush the LDS tile to the table: every word of the tile is written once.
template <bool XOR>
__device__ __forceinline__ void flush_tile(const TilePlan& P, int tile, const TileBox& B, const uint32_t* lds,
                                           uint32_t* __restrict__ table, bool accumulate) {
    for (int i = threadIdx.x; i < P.tile_words; i += kBlock) {
        uint64_t g;
        if (P.kind == TILE_MORTON_CUBE) {
            g = (uint64_t)tile * (uint64_t)P.tile_words + (uint64_t)i;
        } else {
            const int row = i / P.row_words, w = i - row * P.row_words;
            const int y = B.y0 + row % P.ty, z = B.z0 + row / P.ty;
            g = ((uint64_t)y + (uint64_t)z * (uint64_t)P.G) * (uint64_t)P.row_words + (uint64_t)w;
        }
        uint32_t v = lds[i];
        if (accumulate) v = XOR ? (v ^ table[g]) : (v | table[g]);
        table[g] = v;
    }
}

// Work unit: one tile of one job (global tile id = tile_base + local tile).
__global__ void __launch_bounds__(kBlock) raster_surface_kernel(ItemDev one, ItemMap M,
                                                               const unsigned* __restrict__ offsets,
                                                               unsigned* __restrict__ counts,
                                                               const unsigned* __restrict__ list) {
    extern __shared__ uint32_t lds[];
    __shared__ WaveRaster WR[kBlock / 64];
    __shared__ BlockJob J;
    resolve_job(one, M, blockIdx.x, J);
    if (J.local == ~0u) return;
    const vox_info info = J.it.info;
    const TilePlan P = J.it.P;
    const int tile = (int)J.local;
    const unsigned gtile = J.it.tile_base + J.local;
    const uint64_t tri_base = J.it.tri_base;
    const float* __restrict__ tris = J.it.tris;
    const TileBox B = tile_box(P, tile);
    for (int i = threadIdx.x; i < P.tile_words; i += kBlock) lds[i] = 0u;
    if (threadIdx.x == 0) counts[gtile] = 0u;  // leave the workspace zeroed for the next call
    __syncthreads();
    const unsigned beg = offsets[gtile], end = offsets[gtile + 1];
    WaveRaster& W = WR[threadIdx.x >> 6];
    for (unsigned base = beg; base < end; base += kBlock) {
        const unsigned i = base + threadIdx.x;
        const bool have = i < end;
        float v[3][3];
        if (have) {
            load_tri(tris, (uint64_t)list[i] - tri_base, info, v);
        } else {
#pragma unroll
            for (int k = 0; k < 3; k++) v[k][0] = v[k][1] = v[k][2] = 0.0f;
        }
        SurfTri T;
        surface_setup(v, info, T);
        int xa, xb, ya, yb, za, zb;
        clip_box(T, B, xa, xb, ya, yb, za, zb);
        if (P.kind == TILE_MORTON_CUBE) {
            wave_raster(W, info, have, T, xa, xb, ya, yb, za, zb, [&](int y, int z, int cx, uint32_t mask) {
                const uint32_t ly = (uint32_t)(y - B.y0), lz = (uint32_t)(z - B.z0);
                const uint64_t base_local = ((uint64_t)spread10(ly) << 1) | ((uint64_t)spread10(lz) << 2);
                const uint32_t yz_low = ((ly & 1u) << 1) | ((lz & 1u) << 2) | (((ly >> 1) & 1u) << 4);
                morton_chunk_words(base_local, yz_low, mask >> (B.x0 - cx),
                                   [&](uint64_t w, uint32_t bits) { atomicOr(lds + (uint32_t)w, bits); });
            });
        } else {
            wave_raster(W, info, have, T, xa, xb, ya, yb, za, zb, [&](int y, int z, int cx, uint32_t mask) {
                const int row = (y - B.y0) + (z - B.z0) * P.ty;
                atomicOr(lds + row * P.row_words + (cx >> 5), __brev(mask));
            });
        }
    }
    __syncthreads();
    flush_tile<false>(P, tile, B, lds, J.it.table, J.it.accumulate != 0);
}

// ---------------------------------------------------------------------------
// 4b: solid raster (markers + in-row suffix XOR), one workgroup per tile
// ---------------------------------------------------------------------------
// Suffix-XOR every row of the LDS tile in place: bit x becomes the XOR of the
// markers at x' >= x.  Bits are MSB-first (x = 32w + j at bit 31-j), so within
// a word that is a prefix XOR from bit 0 upward; whole words with larger x
// contribute their marker parity.  One lane per word, rows of <= 64 words.
__device__ __forceinline__ void suffix_xor_rows(const TilePlan& P, uint32_t* lds) {
    const int rw = P.row_words;
    const int rows = P.ty * P.tz;
    const int per_wave = 64 / rw;
    const int lane = __lane_id();
    const int wave = threadIdx.x >> 6, nwaves = kBlock / 64;
    const int my_r = lane / rw, my_w = lane - my_r * rw;
    for (int r0 = wave * per_wave; r0 < rows; r0 += nwaves * per_wave) {
        const int r = r0 + my_r;
        const bool act = my_r < per_wave && r < rows;
        uint32_t m = act ? lds[r * rw + my_w] : 0u;
        const unsigned long long par = __ballot(act && (__popc(m) & 1));
        uint32_t s = m;
        s ^= s << 1;
        s ^= s << 2;
        s ^= s << 4;
        s ^= s << 8;
        s ^= s << 16;
        // parity of the words of my row with a larger x (lanes my_w+1 .. rw-1)
        const int row_base = my_r * rw;
        const int hi_cnt = rw - 1 - my_w;
        unsigned long long hi_mask = 0ull;
        if (act && hi_cnt > 0) hi_mask = ((hi_cnt >= 64) ? ~0ull : ((1ull << hi_cnt) - 1ull)) << (row_base + my_w + 1);
        const uint32_t carry = (uint32_t)(__popcll(par & hi_mask) & 1);
        if (act) lds[r * rw + my_w] = s ^ (carry ? 0xFFFFFFFFu : 0u);
    }
}

// Morton output of a row tile: 32-bit words of the G/T aligned cubes covering
// all x of the tile's (y, z) slab, gathered from the linear rows in LDS.
__device__ __forceinline__ void flush_morton_rows(const TilePlan& P, int tile, const TileBox& B, const uint32_t* lds,
                                                  uint32_t* __restrict__ table, bool accumulate) {
    const int T = P.ty;  // cube edge (ty == tz)
    const int cube_words = (T * T * T) >> 5;
    const int ncubes = P.G / T;
    const int cy = B.y0 / T, cz = B.z0 / T;
    const int total = ncubes * cube_words;
    for (int i = threadIdx.x; i < total; i += kBlock) {
        const int cx = i / cube_words, k = i - cx * cube_words;
        // word k of the cube holds local codes 32k .. 32k+31: a 4(x) x 4(y) x 2(z) brick
        const uint32_t code0 = (uint32_t)k << 5;
        uint32_t lx = 0, ly = 0, lz = 0;
        for (int b = 0; b < 10; b++) {
            lx |= ((code0 >> (3 * b)) & 1u) << b;
            ly |= ((code0 >> (3 * b + 1)) & 1u) << b;
            lz |= ((code0 >> (3 * b + 2)) & 1u) << b;
        }
        const int x0 = cx * T + (int)lx;  // multiple of 4
        uint32_t out = 0;
#pragma unroll
        for (int zz = 0; zz < 2; zz++) {
#pragma unroll
            for (int yy = 0; yy < 4; yy++) {
                const int row = ((int)ly + yy) + ((int)lz + zz) * P.ty;
                const uint32_t w = lds[row * P.row_words + (x0 >> 5)];
                const uint32_t nib = (w >> (28 - (x0 & 31))) & 0xFu;  // bit3 = x0, bit0 = x0+3
#pragma unroll
                for (int xx = 0; xx < 4; xx++) {
                    if ((nib >> (3 - xx)) & 1u) {
                        const uint32_t j = (uint32_t)(xx & 1) | ((uint32_t)(yy & 1) << 1) | ((uint32_t)zz << 2) |
                                           ((uint32_t)(xx >> 1) << 3) | ((uint32_t)(yy >> 1) << 4);
                        out |= 1u << (31u - j);
                    }
                }
            }
        }
        const uint64_t g = morton3((uint32_t)cx, (uint32_t)cy, (uint32_t)cz) * (uint64_t)cube_words + (uint64_t)k;
        if (accumulate) out ^= table[g];
        table[g] = out;
    }
}

// Spill records: (first bit, last bit | job << 34) -- bit indices are < 2^33.
constexpr int kSpillJobShift = 34;

__global__ void __launch_bounds__(kBlock) raster_solid_kernel(ItemDev one, ItemMap M,
                                                             const unsigned* __restrict__ offsets,
                                                             unsigned* __restrict__ counts,
                                                             const unsigned* __restrict__ list,
                                                             unsigned* __restrict__ misc,
                                                             unsigned long long* __restrict__ spill,
                                                             unsigned spill_cap) {
    extern __shared__ uint32_t lds[];
    __shared__ BlockJob J;
    resolve_job(one, M, blockIdx.x, J);
    if (J.local == ~0u) return;
    const vox_info info = J.it.info;
    const TilePlan P = J.it.P;
    const int tile = (int)J.local;
    const unsigned gtile = J.it.tile_base + J.local;
    const uint64_t tri_base = J.it.tri_base;
    const float* __restrict__ tris = J.it.tris;
    const TileBox B = tile_box(P, tile);
    const int G = P.G;
    for (int i = threadIdx.x; i < P.tile_words; i += kBlock) lds[i] = 0u;
    if (threadIdx.x == 0) counts[gtile] = 0u;
    __syncthreads();
    unsigned long long events = 0;
    const unsigned beg = offsets[gtile], end = offsets[gtile + 1];
    for (unsigned i = beg + threadIdx.x; i < end; i += kBlock) {
        float v[3][3];
        load_tri(tris, (uint64_t)list[i] - tri_base, info, v);
        SolidTri S;
        solid_setup(v, info, S);
        int ylo, yhi, zlo, zhi;
        solid_columns(S, G, ylo, yhi, zlo, zhi, nullptr);
        ylo = max(ylo, B.y0); yhi = min(yhi, B.y1);
        zlo = max(zlo, B.z0); zhi = min(zhi, B.z1);
        for (int y = ylo; y <= yhi; y++) {
            const float py = ((float)y + 0.5f) * info.unit[1];
            for (int z = zlo; z <= zhi; z++) {
                const float pz = ((float)z + 0.5f) * info.unit[2];
                if (!column_hit(S, py, pz)) continue;
                const int64_t xmax = column_xmax(S, info, py, pz);
                if (xmax < 0) {
                    events++;
                    continue;
                }
                int xm = (int)(xmax < G ? xmax : (int64_t)(G - 1));
                if (xmax >= G) {
                    events++;
                    if (P.kind == TILE_LINEAR) {
                        // bits past the row continue along the linear index
                        const uint64_t row0 = ((uint64_t)y + (uint64_t)z * (uint64_t)G) * (uint64_t)G;
                        const uint64_t total = (uint64_t)G * (uint64_t)G * (uint64_t)G;
                        const uint64_t s0 = row0 + (uint64_t)G;
                        uint64_t s1 = row0 + (uint64_t)xmax;
                        if (s1 > total - 1) s1 = total - 1;
                        if (s0 <= s1) {
                            const unsigned slot = atomicAdd(misc + MISC_SPILL, 1u);
                            if (slot < spill_cap) {
                                spill[2 * slot] = s0;
                                spill[2 * slot + 1] = s1 | ((uint64_t)J.index << kSpillJobShift);
                            } else {
                                misc[MISC_SPILL_OVF] = 1u;
                            }
                        }
                    }
                }
                const int row = (y - B.y0) + (z - B.z0) * P.ty;
                atomicXor(lds + row * P.row_words + (xm >> 5), 1u << (31u - (uint32_t)(xm & 31)));
            }
        }
    }
    if (J.it.offrow && events) atomicAdd(J.it.offrow, events);
    __syncthreads();
    suffix_xor_rows(P, lds);
    __syncthreads();
    if (P.kind == TILE_LINEAR)
        flush_tile<true>(P, tile, B, lds, J.it.table, J.it.accumulate != 0);
    else
        flush_morton_rows(P, tile, B, lds, J.it.table, J.it.accumulate != 0);
}

// ---------------------------------------------------------------------------
// 5: wide triangles (and solid spills), global atomics, after the raster pass
// ---------------------------------------------------------------------------
// Wide triangles: global triangle ids, job found through M.base = per-job
// triangle prefix.  One workgroup per wide triangle (persistent grid); its
// (row, 32-voxel chunk) items are split across the workgroup.

__global__ void __launch_bounds__(kBlock) wide_surface_kernel(ItemDev one, ItemMap M,
                                                             const unsigned* __restrict__ wide_ids,
                                                             const unsigned* __restrict__ misc) {
    const unsigned nw = misc[MISC_WIDE_N];
    for (unsigned w = blockIdx.x; w < nw; w += gridDim.x) {
        const unsigned gt = wide_ids[w];
        const ItemDev& it = job_ref(one, M, job_of_triangle(M, gt));
        const vox_info info = it.info;
        const uint32_t G = info.gridsize[0];
        uint32_t* __restrict__ table = it.table;
        const int layout = it.layout;
        float v[3][3];
        load_tri(it.tris, (uint64_t)gt - it.tri_base, info, v);
        SurfTri T;
        surface_setup(v, info, T);
        const int ny = T.hi[1] - T.lo[1] + 1, nz = T.hi[2] - T.lo[2] + 1;
        const int c0 = T.lo[0] >> 5, nch = (T.hi[0] >> 5) - c0 + 1;
        const uint64_t items = (uint64_t)ny * nz * nch;
        for (uint64_t k = threadIdx.x; k < items; k += kBlock) {
            const int ch = (int)(k % nch);
            const uint64_t r = k / nch;
            const int y = T.lo[1] + (int)(r % ny), z = T.lo[2] + (int)(r / ny);
            const float py = (float)y * info.unit[1], pz = (float)z * info.unit[2];
            if (!edges_pass(T.yz, py, pz)) continue;
            const int cx = (c0 + ch) << 5;
            const int xa = max(T.lo[0], cx), xb = min(T.hi[0], cx + 31);
            uint64_t cur = ~0ull;
            uint32_t mask = 0;
            for (int x = xa; x <= xb; x++) {
                const float px = (float)x * info.unit[0];
                if (!(plane_pass(T, px, py, pz) && edges_pass(T.xy, px, py) && edges_pass(T.zx, pz, px))) continue;
                const uint64_t idx = voxel_index(layout, G, x, y, z);
                if ((idx >> 5) != cur) {
                    if (mask) atomicOr(table + cur, mask);
                    cur = idx >> 5;
                    mask = 0;
                }
                mask |= bit_mask(idx);
            }
            if (mask) atomicOr(table + cur, mask);
        }
    }
}

__global__ void __launch_bounds__(kBlock) wide_solid_kernel(ItemDev one, ItemMap M,
                                                           const unsigned* __restrict__ wide_ids,
                                                           const unsigned* __restrict__ misc,
                                                           const unsigned long long* __restrict__ spill,
                                                           unsigned spill_cap) {
    const unsigned nw = misc[MISC_WIDE_N];
    const unsigned ns = misc[MISC_SPILL];
    const bool ovf = misc[MISC_SPILL_OVF] != 0u;
    // (a) wide triangles: every hit column XORs its whole run with atomics
    for (unsigned w = blockIdx.x; w < nw; w += gridDim.x) {
        const unsigned gt = wide_ids[w];
        const ItemDev& it = job_ref(one, M, job_of_triangle(M, gt));
        const vox_info info = it.info;
        const int G = it.P.G;
        float v[3][3];
        load_tri(it.tris, (uint64_t)gt - it.tri_base, info, v);
        SolidTri S;
        solid_setup(v, info, S);
        int ylo, yhi, zlo, zhi;
        solid_columns(S, G, ylo, yhi, zlo, zhi, nullptr);
        const int ny = yhi - ylo + 1, nz = zhi - zlo + 1;
        const uint64_t cols = (ny > 0 && nz > 0) ? (uint64_t)ny * nz : 0;
        unsigned long long events = 0;
        for (uint64_t c = threadIdx.x; c < cols; c += kBlock) {
            const int y = ylo + (int)(c / nz), z = zlo + (int)(c % nz);
            const float py = ((float)y + 0.5f) * info.unit[1], pz = ((float)z + 0.5f) * info.unit[2];
            if (!column_hit(S, py, pz)) continue;
            const int64_t xmax = column_xmax(S, info, py, pz);
            if (xmax < 0 || xmax >= G) events++;
            if (xmax < 0) continue;
            xor_column_run(it.table, it.layout, G, y, z, xmax);
        }
        if (it.offrow && events) atomicAdd(it.offrow, events);
    }
    // (b) spills recorded by the raster pass (linear layout only); the words
    // of one (possibly long) spilled run are split across the workgroup
    if (!ovf) {
        const unsigned m = min(ns, spill_cap);
        for (unsigned s = blockIdx.x; s < m; s += gridDim.x) {
            const uint64_t b0 = spill[2 * s];
            const uint64_t b1 = spill[2 * s + 1] & ((1ull << kSpillJobShift) - 1ull);
            uint32_t* __restrict__ table = job_ref(one, M, (int)(spill[2 * s + 1] >> kSpillJobShift)).table;
            for (uint64_t w = (b0 >> 5) + threadIdx.x; w <= (b1 >> 5); w += kBlock) {
                const uint32_t lo = (w == (b0 >> 5)) ? (uint32_t)(b0 & 31u) : 0u;
                const uint32_t hi = (w == (b1 >> 5)) ? (uint32_t)(b1 & 31u) : 31u;
                atomicXor(table + w, (0xFFFFFFFFu >> lo) & (0xFFFFFFFFu << (31u - hi)));
            }
        }
    } else {
        // the spill list overflowed: recompute the spills of every binned triangle
        const int nj = M.items ? M.nitems : 1;
        const uint64_t total = M.items ? (uint64_t)M.base[M.nitems] : one.info.n_triangles;
        for (uint64_t gt = (uint64_t)blockIdx.x * kBlock + threadIdx.x; gt < total;
             gt += (uint64_t)gridDim.x * kBlock) {
            const int j = nj > 1 ? job_of_triangle(M, (uint32_t)gt) : 0;
            const ItemDev& it = job_ref(one, M, j);
            if (it.layout != VOX_LAYOUT_LINEAR) continue;
            const vox_info info = it.info;
            const TilePlan P = it.P;
            const int G = P.G;
            float v[3][3];
            load_tri(it.tris, gt - it.tri_base, info, v);
            SolidTri S;
            solid_setup(v, info, S);
            int ylo, yhi, zlo, zhi;
            if (!S.valid || !solid_columns(S, G, ylo, yhi, zlo, zhi, nullptr)) continue;
            const int ntl = (yhi / P.ty - ylo / P.ty + 1) * (zhi / P.tz - zlo / P.tz + 1);
            if (ntl > P.max_tiles_per_tri) continue;  // wide: already applied in full
            for (int y = ylo; y <= yhi; y++) {
                const float py = ((float)y + 0.5f) * info.unit[1];
                for (int z = zlo; z <= zhi; z++) {
                    const float pz = ((float)z + 0.5f) * info.unit[2];
                    if (!column_hit(S, py, pz)) continue;
                    const int64_t xmax = column_xmax(S, info, py, pz);
                    if (xmax < G) continue;
                    const uint64_t tot = (uint64_t)G * (uint64_t)G * (uint64_t)G;
                    const uint64_t row0 = ((uint64_t)y + (uint64_t)z * (uint64_t)G) * (uint64_t)G;
                    uint64_t s1 = row0 + (uint64_t)xmax;
                    if (s1 > tot - 1) s1 = tot - 1;
                    if (row0 + G <= s1) xor_linear_run(it.table, row0 + G, s1);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// host side: plans, workspace layout, launches
// ---------------------------------------------------------------------------
static int ilog2(int v) {
    int r = 0;
    while ((1 << (r + 1)) <= v) r++;
    return r;
}

bool make_plan(int G, int layout, bool solid, TilePlan& P, int scale) {
    P = TilePlan{};
    P.G = G;
    P.max_tiles_per_tri = 8;
    const uint64_t cells = (uint64_t)G * G * G;
    if (layout == VOX_LAYOUT_LINEAR) {
        if (G % 32 != 0 || G > 2048) return false;
        // surface tiles ~4 KB (>= 2048 tiles when the grid allows); solid
        // tiles larger (its per-tile work is one column loop per triangle)
        uint64_t bits = solid ? cells / 4096 : cells / 2048;
        const uint64_t lo = solid ? 32768 : 1024, hi = solid ? 262144 : 32768;
        if (bits < lo) bits = lo;
        if (bits > hi) bits = hi;
        if (bits < cells / 65536) bits = cells / 65536;  // at most 65536 tiles
        // tuned size: half or twice the default tile (within the limits below)
        if (scale < 0 && bits > lo) bits /= 2;
        if (scale > 0 && bits * 2 <= hi && bits * 2 <= 524288) bits *= 2;
        int rows = (int)(bits / (uint64_t)G);
        if (rows < 1) rows = 1;
        rows = 1 << ilog2(rows);
        if (rows > G * G) rows = G * G;
        const int lg = ilog2(rows);
        P.ty = 1 << ((lg + 1) / 2);
        P.tz = rows / P.ty;
        if (P.ty > G) { P.tz *= P.ty / G; P.ty = G; }
        if (P.tz > G) P.tz = G;
        P.kind = TILE_LINEAR;
        P.tx = G;
        P.ntx = 1;
        P.nty = G / P.ty;
        P.ntz = G / P.tz;
        P.row_words = G / 32;
        P.tile_words = P.row_words * P.ty * P.tz;
    } else {
        if ((G & (G - 1)) != 0) return false;
        if (!solid) {
            if (G < 4) return false;
            int T = G <= 64 ? (G < 8 ? G : 8) : (G <= 256 ? 16 : 32);
            if (scale < 0 && T >= 8) T /= 2;
            if (scale > 0 && T * 2 <= G && T * 2 <= 32) T *= 2;
            P.kind = TILE_MORTON_CUBE;
            P.tx = P.ty = P.tz = T;
            P.ntx = P.nty = P.ntz = G / T;
            P.tile_words = (T * T * T) >> 5;
            P.lg_cube = ilog2(T);
            P.row_words = 0;
        } else {
            if (G < 32 || G > 2048) return false;  // rows narrower than a word: direct path
            int T = G >= 512 ? 16 : 8;
            if (scale < 0 && T >= 8) T /= 2;
            if (scale > 0 && T * 2 <= 16 && T * 2 <= G) T *= 2;
            P.kind = TILE_MORTON_ROWS;
            P.tx = G;
            P.ty = P.tz = T;
            P.ntx = 1;
            P.nty = P.ntz = G / T;
            P.row_words = G / 32;
            P.tile_words = P.row_words * T * T;
            P.lg_cube = ilog2(T);
        }
    }
    P.ntiles = P.ntx * P.nty * P.ntz;
    if (P.ntiles > 65536 || P.tile_words * 4 > 65536) return false;
    return true;
}

// Fixed-order layout for capacities (tiles_cap, n_cap): misc counters first,
// then the per-tile counts -- both must be zero before a call and are left
// zero by it -- then offsets, lists, wide ids and spill records.
WorkspaceLayout workspace_layout(uint64_t tiles_cap, uint64_t n_cap, int max_tiles_per_tri) {
    WorkspaceLayout L{};
    auto align = [](uint64_t v) { return (v + 255) & ~255ull; };
    L.list_cap = n_cap * (uint64_t)max_tiles_per_tri;
    L.wide_cap = n_cap;
    L.spill_cap = 65536;
    uint64_t off = 0;
    L.misc_off = off; off = align(off + 4ull * MISC_WORDS);
    L.counts_off = off; off = align(off + 4ull * (tiles_cap + 1));
    L.offsets_off = off; off = align(off + 4ull * (tiles_cap + 1));
    L.list_off = off; off = align(off + 4ull * (L.list_cap > 0 ? L.list_cap : 1));
    L.wide_off = off; off = align(off + 4ull * (L.wide_cap > 0 ? L.wide_cap : 1));
    L.spill_off = off; off = align(off + 16ull * L.spill_cap);
    L.total = off;
    return L;
}

// Launch the tile pipeline for one job (items == nullptr) or a job array.
// `M_bin`: base = per-job bin blocks; `M_tile`: base = per-job tiles;
// `M_tri`: base = per-job triangle prefix.  total_* are the launch sizes.
hipError_t launch_tiled_jobs(const ItemDev& one, const ItemMap& M_bin, const ItemMap& M_tile, const ItemMap& M_tri,
                             uint32_t total_bin_blocks, uint32_t total_tiles, bool solid, uint32_t max_tile_words,
                             void* ws, const WorkspaceLayout& L, hipStream_t stream) {
    char* wb = (char*)ws;
    unsigned* counts = (unsigned*)(wb + L.counts_off);
    unsigned* offsets = (unsigned*)(wb + L.offsets_off);
    unsigned* misc = (unsigned*)(wb + L.misc_off);
    unsigned* list = (unsigned*)(wb + L.list_off);
    unsigned* wide = (unsigned*)(wb + L.wide_off);
    unsigned long long* spill = (unsigned long long*)(wb + L.spill_off);
    const unsigned nb = total_bin_blocks > 0 ? total_bin_blocks : 1u;
    const unsigned ntl = total_tiles > 0 ? total_tiles : 1u;
    const size_t lds = (size_t)max_tile_words * 4;
    if (solid) {
        bin_kernel<true, false><<<nb, kBlock, 0, stream>>>(one, M_bin, counts, list, misc, wide);
        bin_scan_kernel<<<1, 1024, 0, stream>>>(total_tiles, counts, offsets, misc);
        bin_kernel<true, true><<<nb, kBlock, 0, stream>>>(one, M_bin, counts, list, misc, wide);
        raster_solid_kernel<<<ntl, kBlock, lds, stream>>>(one, M_tile, offsets, counts, list, misc, spill,
                                                          (unsigned)L.spill_cap);
        wide_solid_kernel<<<kWideBlocks, kBlock, 0, stream>>>(one, M_tri, wide, misc, spill, (unsigned)L.spill_cap);
    } else {
        bin_kernel<false, false><<<nb, kBlock, 0, stream>>>(one, M_bin, counts, list, misc, wide);
        bin_scan_kernel<<<1, 1024, 0, stream>>>(total_tiles, counts, offsets, misc);
        bin_kernel<false, true><<<nb, kBlock, 0, stream>>>(one, M_bin, counts, list, misc, wide);
        raster_surface_kernel<<<ntl, kBlock, lds, stream>>>(one, M_tile, offsets, counts, list);
        wide_surface_kernel<<<kWideBlocks, kBlock, 0, stream>>>(one, M_tri, wide, misc);
    }
    return hipGetLastError();
}

hipError_t launch_tiled(const vox_info& info, const float* tris, uint32_t* table, const TilePlan& P, bool solid,
                        int layout, bool accumulate, void* ws, const WorkspaceLayout& L,
                        unsigned long long* offrow, hipStream_t stream) {
    ItemDev one{};
    one.info = info;
    one.tris = tris;
    one.table = table;
    one.offrow = offrow;
    one.P = P;
    one.layout = layout;
    one.solid = solid ? 1 : 0;
    one.accumulate = accumulate ? 1 : 0;
    one.tile_base = 0;
    one.tri_base = 0;
    const ItemMap none{nullptr, nullptr, 1, 0};
    const uint64_t blocks = (info.n_triangles + kBlock - 1) / kBlock;
    return launch_tiled_jobs(one, none, none, none, (uint32_t)blocks, (uint32_t)P.ntiles, solid,
                             (uint32_t)P.tile_words, ws, L, stream);
}

}  // namespace voxamd
