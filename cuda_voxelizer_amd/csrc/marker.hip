// marker.hip -- solid voxelization by run-end markers and a row suffix XOR.
//
// The reference XORs, for every hit column (y, z) of every triangle, the
// whole run x = 0..xmax into the table (cpu_voxelizer.cpp:294-305): O(xmax)
// bit writes per crossing.  The final bit (x, y, z) is the parity of the runs
// of its column with xmax >= x, so it is enough to toggle ONE marker bit at
// (xmax, y, z) per run and then replace every row by its suffix XOR along x:
//
//   1. fill      zero the marker table: the output table itself (linear
//                layout, no accumulate) or a linear scratch table
//   2. markers   per triangle (one lane each; triangles with many columns
//                are split over the workgroup): solid setup, column tests,
//                xmax, one non-returning atomicXor per hit column
//   3. suffix    per row (G/32 words, lanes = words): in-word prefix XOR from
//                bit 0 up (larger x = lower bit) plus the parity of the words
//                of larger x (a ballot).  In place for the linear layout;
//                from the scratch table, one workgroup per block of 32 x rz
//                rows staged in LDS, for the Morton layout (4x4x2 brick words
//                assembled from 8 rows at a time) and for accumulation
//                (XOR into the caller's table)
//   4. spills    runs with xmax >= G continue along the linear index into the
//                following rows (as the reference indexes them): their
//                in-row part is a marker at G-1, the rest is recorded and
//                XORed into the final table afterwards (word masks, atomics)
//
// G % 32 == 0 (rows are whole words); Morton: power-of-two G up to 4096.
// Per-column arithmetic: vox_device.h (the reference CPU path bit for bit).
#include "bricks.h"
#include "jobs.h"
#include "region.h"
#include "vox_device.h"
#include "vox_host.h"
#include "vox_kernels.h"

#include <map>
#include <mutex>
#include <utility>

#pragma clang fp contract(off)

namespace voxamd {

constexpr int kMBlock = 256;
constexpr int kMQueue = kMBlock;   // triangles with many columns, split over the workgroup
constexpr int kMBigCols = 32;      // columns above which a triangle is queued (the bunny at 64^3:
                                   // 0.0325 vs 0.0349 ms at 64; the benchmark spheres unchanged)
constexpr uint32_t kSpillCap = 1u << 16;
constexpr int kJobShift = 42;  // spill records: job index above the bit index (G^3 <= 2^42)
constexpr unsigned long long kBitMask = (1ull << kJobShift) - 1ull;
constexpr uint32_t kSuffixRows = 256;  // rows per work unit of the job-mapped suffix
constexpr int kSpillBlocks = 64;      // spill pass workgroups (each takes whole records)

// misc words: [0] spill count, [1] spill overflow flag, [2] spill blocks done.
// Zero when allocated; the spill pass returns them to zero, so every call
// starts from zero without a reset launch (and the Morton layout, which
// records no spills, never touches them).
struct MarkerWs {
    unsigned* misc;
    unsigned long long* spill;  // [2*i] job << kJobShift | first bit, [2*i+1] last bit (inclusive)
};

// Region algorithm: the workgroup's markers, merged by XOR in an LDS hash
// keyed in the region key space, then flushed as region pairs.
constexpr int kMHashBits = 11;
constexpr int kMHash = 1 << kMHashBits;
template <bool REG>
struct MarkerLds {};
template <>
struct MarkerLds<true> {
    uint32_t keys[kMHash];
    uint32_t vals[kMHash];
    RegionFlushShared flush;
};

__device__ __forceinline__ void marker_xor(MarkerLds<true>& H, uint32_t key, uint32_t mask, const RegionArgs& RA) {
    uint32_t h = (key * 2654435761u) >> (32 - kMHashBits);
#pragma unroll 1
    for (int probe = 0; probe < kRegionProbes; probe++) {
        const uint32_t k = atomicCAS(&H.keys[h], kRegionEmpty, key);  // claims or finds, in one round trip
        if (k == kRegionEmpty || k == key) {
            atomicXor(&H.vals[h], mask);
            return;
        }
        h = (h + 1) & (kMHash - 1);
    }
    region_fallback(RA.P, RA.W, key, mask);
}

// One hit column: marker toggle (and spill record) of the run x = 0..xmax.
// REG: the marker goes to the LDS hash H instead of the table.
template <bool REG>
__device__ __forceinline__ void column_run(uint32_t* __restrict__ table, int64_t xmax, int y, int z, uint32_t G,
                                           uint64_t total_bits, const MarkerWs& ws, unsigned long long& offrow,
                                           bool spills, unsigned long long jobtag, MarkerLds<REG>& H,
                                           const RegionArgs& RA) {
    if (xmax < 0 || xmax >= (int64_t)G) offrow++;
    if (xmax < 0) return;
    const uint64_t row = (uint64_t)y * G + (uint64_t)z * G * G;
    const uint32_t xm = xmax < (int64_t)G ? (uint32_t)xmax : G - 1;
    if constexpr (REG) {
        marker_xor(H, region_row_key(RA.P, xm >> 5, (uint32_t)y, (uint32_t)z), 1u << (31u - (xm & 31u)), RA);
    } else {
        const uint64_t idx = row + xm;
        atomicXor(table + (idx >> 5), bit_mask(idx));
    }
    if (spills && xmax >= (int64_t)G && row + G < total_bits) {
        // bits row+G .. row+xmax, clipped to the table
        const uint64_t last = (uint64_t)xmax < total_bits - 1 - row ? row + (uint64_t)xmax : total_bits - 1;
        const unsigned slot = atomicAdd(ws.misc, 1u);
        if (slot < kSpillCap) {
            ws.spill[2 * slot] = jobtag | (row + G);
            ws.spill[2 * slot + 1] = last;
        } else {
            ws.misc[1] = 1u;
        }
    }
}

// Work unit: kMBlock consecutive triangles of one job; the markers go to the
// job's table.  MULTI: the job comes from the launch's job map (block
// prefix); otherwise it is `one`.  REG: the markers are staged as region
// pairs (RA) instead (phase 1 of the region algorithm).
template <bool MULTI, bool REG>
__global__ void __launch_bounds__(kMBlock)
    solid_marker_kernel(ItemDev one, ItemMap M, MarkerWs ws, int spills, RegionArgs RA) {
    __shared__ unsigned queue[kMQueue];
    __shared__ unsigned qn;
    __shared__ BlockJob J;
    __shared__ MarkerLds<REG> H;
    uint32_t local = blockIdx.x;
    int job = 0;
    if (MULTI) {
        resolve_job(one, M, blockIdx.x, J);
        local = J.local;
        job = J.index;
        if (local == ~0u) return;
    }
    const ItemDev& it = MULTI ? J.it : one;
    const vox_info& info = it.info;
    const float* __restrict__ tris = it.tris;
    uint32_t* __restrict__ table = it.table;
    unsigned long long* __restrict__ offrow_out = it.offrow;
    const unsigned long long jobtag = (unsigned long long)job << kJobShift;
    const uint64_t n = info.n_triangles;
    const uint64_t first = (uint64_t)local * kMBlock;
    if (first >= n) return;
    const int count = (int)((n - first) < (uint64_t)kMBlock ? (n - first) : (uint64_t)kMBlock);
    const int G = (int)info.gridsize[0];
    const uint64_t total_bits = (uint64_t)G * G * G;
    if (threadIdx.x == 0) qn = 0;
    if constexpr (REG) {
        for (int i = threadIdx.x; i < kMHash; i += kMBlock) {
            H.keys[i] = kRegionEmpty;
            H.vals[i] = 0u;
        }
    }
    __syncthreads();  // qn (and the LDS hash) set
    unsigned long long offrow = 0;
    if ((int)threadIdx.x < count) {
        float v[3][3];
        load_tri(tris, first + threadIdx.x, info, v);
        SolidTri S;
        solid_setup(v, info, S);
        if (S.valid) {
            // column range, clamped to the grid, and its out-of-grid events
            // (as the reference loops from the unclamped bounds)
            int ylo = (int)S.ymin, zlo = (int)S.zmin;
            if (ylo < 0) { offrow++; ylo = 0; }
            if (zlo < 0) { offrow++; zlo = 0; }
            if (S.ymax > (float)(G - 1)) offrow++;
            if (S.zmax > (float)(G - 1)) offrow++;
            const int yhi = S.ymax >= (float)(G - 1) ? G - 1 : (S.ymax >= 0.0f ? (int)S.ymax : -1);
            const int zhi = S.zmax >= (float)(G - 1) ? G - 1 : (S.zmax >= 0.0f ? (int)S.zmax : -1);
            const int64_t cols = (yhi >= ylo && zhi >= zlo) ? (int64_t)(yhi - ylo + 1) * (zhi - zlo + 1) : 0;
            bool queued = false;
            if (cols > (it.big ? (int64_t)it.big : (int64_t)kMBigCols)) {
                const unsigned q = atomicAdd(&qn, 1u);
                if (q < (unsigned)kMQueue) {
                    queue[q] = threadIdx.x;
                    queued = true;
                }
            }
            if (!queued) {
                for (int y = ylo; y <= yhi; y++) {
                    const float py = ((float)y + 0.5f) * info.unit[1];
                    for (int z = zlo; z <= zhi; z++) {
                        const float pz = ((float)z + 0.5f) * info.unit[2];
                        if (!column_hit(S, py, pz)) continue;
                        column_run<REG>(table, column_xmax(S, info, py, pz), y, z, (uint32_t)G, total_bits, ws,
                                        offrow, spills != 0, jobtag, H, RA);
                    }
                }
            }
        }
    }
    __syncthreads();
    const unsigned nq = qn < (unsigned)kMQueue ? qn : (unsigned)kMQueue;
    for (unsigned q = 0; q < nq; q++) {
        const int lt = (int)queue[q];
        float v[3][3];
        load_tri(tris, first + (uint64_t)lt, info, v);  // the same triangle in every lane
        SolidTri S;
        solid_setup(v, info, S);
        const int ylo = (int)S.ymin < 0 ? 0 : (int)S.ymin, zlo = (int)S.zmin < 0 ? 0 : (int)S.zmin;
        const int yhi = S.ymax >= (float)(G - 1) ? G - 1 : (S.ymax >= 0.0f ? (int)S.ymax : -1);
        const int zhi = S.zmax >= (float)(G - 1) ? G - 1 : (S.zmax >= 0.0f ? (int)S.zmax : -1);
        const int ny = yhi - ylo + 1, nz = zhi - zlo + 1;
        const int64_t cols = (int64_t)ny * nz;
        for (int64_t c = threadIdx.x; c < cols; c += kMBlock) {
            const int y = ylo + (int)(c / nz), z = zlo + (int)(c % nz);
            const float py = ((float)y + 0.5f) * info.unit[1];
            const float pz = ((float)z + 0.5f) * info.unit[2];
            if (!column_hit(S, py, pz)) continue;
            column_run<REG>(table, column_xmax(S, info, py, pz), y, z, (uint32_t)G, total_bits, ws, offrow,
                            spills != 0, jobtag, H, RA);
        }
    }
    if (offrow_out && offrow) atomicAdd(offrow_out, offrow);
    if constexpr (REG) {
        __syncthreads();  // every marker of the workgroup is in the hash
        constexpr int kPer = kMHash / kMBlock;
        uint32_t kk[kPer], vv[kPer];
#pragma unroll
        for (int u = 0; u < kPer; u++) {
            kk[u] = H.keys[threadIdx.x + u * kMBlock];
            vv[u] = H.vals[threadIdx.x + u * kMBlock];
        }
        region_flush<kMBlock, kPer>(H.flush, kk, vv, H.keys, H.vals, RA.P, RA.W);
    }
}

// Rows r_begin..r_end-1 of rw = G/32 words, in place; waves `wave` of
// `nwaves` stride over them.  rw <= 64: a wave holds 64/rw whole rows (lane =
// word); rw > 64: a wave walks its row from the last 64-word segment down,
// carrying the parity.
__device__ __forceinline__ void suffix_rows_range(uint32_t* __restrict__ table, uint32_t G, uint64_t r_begin,
                                                  uint64_t r_end, uint64_t wave, uint64_t nwaves) {
    const uint32_t rw = G >> 5;
    const int lane = __lane_id();
    if (rw <= 64) {
        const uint32_t per = 64 / rw;             // rows per wave step
        const uint32_t r_in = (uint32_t)lane / rw, j = (uint32_t)lane % rw;
        const bool active_lane = r_in < per;
        for (uint64_t r0 = r_begin + wave * per; r0 < r_end; r0 += nwaves * per) {
            const uint64_t r = r0 + r_in;
            const bool act = active_lane && r < r_end;
            const uint64_t idx = r * rw + j;
            const uint32_t w = act ? table[idx] : 0u;
            const unsigned long long par = __ballot(act && (__popc(w) & 1));
            // parity of the words with larger x in the same row: lanes j+1 .. rw-1 of this row
            const uint32_t seg0 = r_in * rw;
            const unsigned long long above = (active_lane && j + 1 < rw)
                                                 ? ((par >> (seg0 + j + 1)) & ((1ull << (rw - 1 - j)) - 1ull))
                                                 : 0ull;
            const uint32_t carry = (uint32_t)(__popcll(above) & 1);
            if (act) table[idx] = prefix_xor_up(w) ^ (carry ? 0xFFFFFFFFu : 0u);
        }
    } else {
        for (uint64_t r = r_begin + wave; r < r_end; r += nwaves) {
            uint32_t carry = 0;  // parity of the words of larger x already done
            for (int64_t hi = (int64_t)rw; hi > 0; hi -= 64) {
                const int64_t lo = hi > 64 ? hi - 64 : 0;
                const bool act = (int64_t)lane < hi - lo;
                const uint64_t idx = r * rw + (uint64_t)lo + (uint64_t)lane;
                const uint32_t w = act ? table[idx] : 0u;
                const unsigned long long par = __ballot(act && (__popc(w) & 1));
                const unsigned long long above = lane < 63 ? (par >> (lane + 1)) : 0ull;
                const uint32_t c = (uint32_t)((__popcll(above) + carry) & 1);
                if (act) table[idx] = prefix_xor_up(w) ^ (c ? 0xFFFFFFFFu : 0u);
                carry = (uint32_t)((__popcll(par) + carry) & 1);
            }
        }
    }
}

// Where the 16-byte row form below applies: G a multiple of 128 and at most
// 64 * 128 (whole rows per wave), the table 16-byte aligned.
__host__ __device__ inline bool rows4_ok(uint32_t G, const uint32_t* table) {
    return (G & 127u) == 0 && G <= 8192u && (((uintptr_t)table) & 15u) == 0;
}

// The same for rows4_ok tables: a lane holds 4 consecutive words (one 16-byte
// load), a row is rw/4 <= 64 lanes, a wave holds 256/rw rows.
// CLEAR: the markers are read from `src` (the library's scratch marker table)
// and the rows written to `table` whole; src words that held markers are
// returned to zero, so the scratch table needs no fill before the next call.
template <bool CLEAR = false>
__device__ __forceinline__ void suffix_rows4_range(uint32_t* __restrict__ table, uint32_t G, uint64_t r_begin,
                                                   uint64_t r_end, uint64_t wave, uint64_t nwaves,
                                                   uint32_t* __restrict__ src = nullptr) {
    const uint32_t lpr = G >> 7;  // lanes per row
    const uint32_t per = 64 / lpr;
    const int lane = __lane_id();
    const uint32_t r_in = (uint32_t)lane / lpr, j = (uint32_t)lane % lpr;
    const bool active_lane = r_in < per;
    uint4* t4 = (uint4*)table;
    uint4* s4 = CLEAR ? (uint4*)src : t4;
    for (uint64_t r0 = r_begin + wave * per; r0 < r_end; r0 += nwaves * per) {
        const uint64_t r = r0 + r_in;
        const bool act = active_lane && r < r_end;
        const uint64_t idx = r * lpr + j;
        uint4 v = act ? s4[idx] : make_uint4(0u, 0u, 0u, 0u);
        if (CLEAR && (v.x | v.y | v.z | v.w)) s4[idx] = make_uint4(0u, 0u, 0u, 0u);
        const uint32_t pa = __popc(v.x) & 1, pb = __popc(v.y) & 1, pc = __popc(v.z) & 1, pd = __popc(v.w) & 1;
        const unsigned long long par = __ballot(act && (pa ^ pb ^ pc ^ pd));
        const uint32_t seg0 = r_in * lpr;
        const unsigned long long above =
            (active_lane && j + 1 < lpr) ? ((par >> (seg0 + j + 1)) & ((1ull << (lpr - 1 - j)) - 1ull)) : 0ull;
        const uint32_t cw = (uint32_t)(__popcll(above) & 1);  // parity of the lanes of larger x
        const uint32_t cz = cw ^ pd, cy = cz ^ pc, cx = cy ^ pb;
        v.w = prefix_xor_up(v.w) ^ (0u - cw);
        v.z = prefix_xor_up(v.z) ^ (0u - cz);
        v.y = prefix_xor_up(v.y) ^ (0u - cy);
        v.x = prefix_xor_up(v.x) ^ (0u - cx);
        if (act) t4[idx] = v;
    }
}

__global__ void __launch_bounds__(256) suffix_rows_kernel(uint32_t* __restrict__ table, uint32_t G,
                                                          uint64_t nrows) {
    const uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    suffix_rows_range(table, G, 0, nrows, wave, nwaves);
}

__global__ void __launch_bounds__(256) suffix_rows4_kernel(uint32_t* __restrict__ table, uint32_t G,
                                                           uint64_t nrows) {
    const uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    suffix_rows4_range(table, G, 0, nrows, wave, nwaves);
}

// Markers in the scratch table -> the rows of the linear table (every word
// written), the scratch returned to zero.
__global__ void __launch_bounds__(256) suffix_scratch_rows4_kernel(uint32_t* __restrict__ markers,
                                                                   uint32_t* __restrict__ table, uint32_t G,
                                                                   uint64_t nrows) {
    const uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    suffix_rows4_range<true>(table, G, 0, nrows, wave, nwaves, markers);
}

// Batched form: work unit = kSuffixRows rows of one job's (linear, in-place)
// table.
__global__ void __launch_bounds__(256) suffix_jobs_kernel(ItemMap M) {
    __shared__ BlockJob J;
    const ItemDev none{};
    resolve_job(none, M, blockIdx.x, J);
    if (J.local == ~0u) return;
    const uint32_t G = J.it.info.gridsize[0];
    uint32_t* __restrict__ table = J.it.table;
    const uint64_t nrows = (uint64_t)G * G;
    const uint64_t r0 = (uint64_t)J.local * kSuffixRows;
    const uint64_t r1 = r0 + kSuffixRows < nrows ? r0 + kSuffixRows : nrows;
    const uint64_t wave = threadIdx.x >> 6;
    if (rows4_ok(G, table))
        suffix_rows4_range(table, G, r0, r1, wave, 4);
    else
        suffix_rows_range(table, G, r0, r1, wave, 4);
}

// Rows of the linear marker table -> output table, one workgroup per block
// of 32 rows in y x rz rows in z (all x), staged in LDS.  Linear output: the
// final words are XORed into (accumulate) or stored to the block's
// contiguous rows.  Morton output: a word is a 4x4x2 brick (code bits 0..4 =
// x0 y0 z0 x1 y1); a thread takes one linear word (32 x) of the 8 rows of a
// brick row and emits the 8 Morton words of its 8 groups of 4 x.
// SUFFIX = false: the rows are copied as they are (linear -> Morton layout
// conversion of a finished table).
// SUFFIX = true: `markers` is the library's scratch marker table, and the block
// returns its rows to zero once loaded, so the next call needs no zero-fill.
template <bool SUFFIX>
__global__ void __launch_bounds__(256) marker_build_kernel(const uint32_t* markers, uint32_t* __restrict__ table,
                                                           uint32_t G, uint32_t rz, int layout, int accumulate) {
    extern __shared__ uint32_t tile[];  // [rz * 32 rows][rw + 1]: the pad word spreads rows over the banks
    const uint32_t rw = G >> 5, nby = G >> 5;
    const uint32_t ts = rw + 1;  // row stride in LDS
    const uint32_t rb = blockIdx.x;
    const uint32_t nrows = 32u * rz;
    const uint32_t z0 = (rb / nby) * rz, y0 = (rb % nby) * 32u;
    // rows (y0.., z) for each z of the block are 32 * rw contiguous words; eight
    // loads in flight per thread before their LDS stores (16-byte loads when
    // rows are whole quads: 32 KB of a 32 x 8 x 1024 block in flight at once)
    const uint32_t total = nrows * rw;
    const bool linear = layout == VOX_LAYOUT_LINEAR;
    // Morton output (power-of-two G) with rows of at most 64 quads: the suffix
    // is taken on the loaded quads (a row's rw / 4 quads sit in consecutive
    // lanes of one wave, and every wave runs the same number of iterations)
    const bool quads = (rw & 3u) == 0 && (((uintptr_t)markers) & 15u) == 0;
    const bool fused = SUFFIX && !linear && quads && rw <= 256u;
    if (quads) {
        const uint32_t total4 = total >> 2, rq = rw >> 2, zq = 8u * rw;  // quads: all, per row, per z
        const uint32_t lane = (uint32_t)__lane_id(), j = lane % rq, seg0 = lane - j;
        const uint4* __restrict__ m4 = (const uint4*)markers;
        for (uint32_t i0 = threadIdx.x; i0 < total4; i0 += 256u * 8u) {
            uint4 v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const uint32_t i = i0 + 256u * (uint32_t)u;
                const uint32_t zl = i / zq, k = i - zl * zq;
                v[u] = i < total4 ? m4[((((uint64_t)(z0 + zl) * G + y0) * rw) >> 2) + k] : make_uint4(0u, 0u, 0u, 0u);
            }
            if (SUFFIX) {
                uint4* c4 = const_cast<uint4*>(m4);
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const uint32_t i = i0 + 256u * (uint32_t)u;
                    const uint32_t zl = i / zq, k = i - zl * zq;
                    // words already zero (rows without markers) need no store
                    if (i < total4 && (v[u].x | v[u].y | v[u].z | v[u].w))
                        c4[((((uint64_t)(z0 + zl) * G + y0) * rw) >> 2) + k] = make_uint4(0u, 0u, 0u, 0u);
                }
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const uint32_t i = i0 + 256u * (uint32_t)u;
                if (fused) {
                    // as suffix_rows4_range: carry = parity of the row's quads of larger x
                    uint4& q = v[u];
                    const uint32_t pa = __popc(q.x) & 1, pb = __popc(q.y) & 1, pc = __popc(q.z) & 1,
                                   pd = __popc(q.w) & 1;
                    const unsigned long long par = __ballot(i < total4 && (pa ^ pb ^ pc ^ pd));
                    const unsigned long long above =
                        j + 1 < rq ? ((par >> (seg0 + j + 1)) & ((1ull << (rq - 1 - j)) - 1ull)) : 0ull;
                    const uint32_t cw = (uint32_t)(__popcll(above) & 1), cz = cw ^ pd, cy = cz ^ pc, cx = cy ^ pb;
                    q.w = prefix_xor_up(q.w) ^ (0u - cw);
                    q.z = prefix_xor_up(q.z) ^ (0u - cz);
                    q.y = prefix_xor_up(q.y) ^ (0u - cy);
                    q.x = prefix_xor_up(q.x) ^ (0u - cx);
                }
                if (i < total4) {
                    uint32_t* d = tile + (i / rq) * ts + (i % rq) * 4u;
                    d[0] = v[u].x;
                    d[1] = v[u].y;
                    d[2] = v[u].z;
                    d[3] = v[u].w;
                }
            }
        }
    } else {
        for (uint32_t i0 = threadIdx.x; i0 < total; i0 += 256u * 8u) {
            uint32_t v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const uint32_t i = i0 + 256u * (uint32_t)u;
                const uint32_t zl = i / (32u * rw), k = i % (32u * rw);
                v[u] = i < total ? markers[((uint64_t)(z0 + zl) * G + y0) * rw + k] : 0u;
            }
            if (SUFFIX) {
                uint32_t* c = const_cast<uint32_t*>(markers);
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const uint32_t i = i0 + 256u * (uint32_t)u;
                    const uint32_t zl = i / (32u * rw), k = i % (32u * rw);
                    if (i < total && v[u]) c[((uint64_t)(z0 + zl) * G + y0) * rw + k] = 0u;
                }
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const uint32_t i = i0 + 256u * (uint32_t)u;
                if (i < total) tile[(i / rw) * ts + i % rw] = v[u];
            }
        }
    }
    __syncthreads();
    if (SUFFIX && !fused) {
        tile_suffix(tile, ts, rw, nrows, [&](uint32_t r, uint32_t j, uint32_t fin) {
            if (linear) {
                const uint64_t idx = ((uint64_t)(z0 + r / 32u) * G + y0 + (r & 31u)) * rw + j;
                table[idx] = accumulate ? (table[idx] ^ fin) : fin;
            } else {
                tile[r * ts + j] = fin;
            }
        });
    }
    if (linear) return;
    __syncthreads();
    tile_to_bricks(tile, ts, table, G, rw, rz, y0, z0, accumulate);
}

// Spilled run parts, XORed into the final tables: from the records (each
// names its job of `M_jobs`, or `one`), one workgroup per record -- or, when
// more than kSpillCap runs spilled, recomputed from every triangle (all of
// `one`, or of every job of `M_tris`, whose base is the triangle prefix).
// The last workgroup to finish returns the counters to zero.
__global__ void __launch_bounds__(256) spill_kernel(ItemDev one, ItemMap M_jobs, ItemMap M_tris, MarkerWs ws) {
    const unsigned spilled = ws.misc[0], overflow = ws.misc[1];
    if (!overflow) {
        const unsigned ns = spilled < kSpillCap ? spilled : kSpillCap;
        for (unsigned s = blockIdx.x; s < ns; s += gridDim.x) {
            const unsigned long long rec = ws.spill[2 * s];
            const uint64_t b0 = rec & kBitMask, b1 = ws.spill[2 * s + 1];
            uint32_t* __restrict__ table = job_ref(one, M_jobs, (int)(rec >> kJobShift)).table;
            for (uint64_t w = (b0 >> 5) + threadIdx.x; w <= (b1 >> 5); w += blockDim.x) {
                const uint32_t lo = w == (b0 >> 5) ? (uint32_t)(b0 & 31u) : 0u;
                const uint32_t hi = w == (b1 >> 5) ? (uint32_t)(b1 & 31u) : 31u;
                atomicXor(table + w, (0xFFFFFFFFu >> lo) & (0xFFFFFFFFu << (31u - hi)));
            }
        }
    } else {
        const uint64_t total = M_tris.items ? (uint64_t)M_tris.base[M_tris.nitems] : one.info.n_triangles;
        for (uint64_t gt = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; gt < total;
             gt += (uint64_t)gridDim.x * blockDim.x) {
            const int j = M_tris.items ? job_of_triangle(M_tris, (uint32_t)gt) : 0;
            const ItemDev& it = job_ref(one, M_tris, j);
            const vox_info& info = it.info;
            uint32_t* __restrict__ table = it.table;
            const uint64_t t = M_tris.items ? gt - M_tris.base[j] : gt;
            const int G = (int)info.gridsize[0];
            const uint64_t total_bits = (uint64_t)G * G * G;
            float v[3][3];
            load_tri(it.tris, t, info, v);
            SolidTri S;
            solid_setup(v, info, S);
            if (!S.valid) continue;
            const int ylo = (int)S.ymin < 0 ? 0 : (int)S.ymin, zlo = (int)S.zmin < 0 ? 0 : (int)S.zmin;
            const int yhi = S.ymax >= (float)(G - 1) ? G - 1 : (S.ymax >= 0.0f ? (int)S.ymax : -1);
            const int zhi = S.zmax >= (float)(G - 1) ? G - 1 : (S.zmax >= 0.0f ? (int)S.zmax : -1);
            for (int y = ylo; y <= yhi; y++) {
                const float py = ((float)y + 0.5f) * info.unit[1];
                for (int z = zlo; z <= zhi; z++) {
                    const float pz = ((float)z + 0.5f) * info.unit[2];
                    if (!column_hit(S, py, pz)) continue;
                    const int64_t xmax = column_xmax(S, info, py, pz);
                    const uint64_t row = (uint64_t)y * G + (uint64_t)z * G * G;
                    if (xmax < (int64_t)G || row + G >= total_bits) continue;
                    const uint64_t b0 = row + G;
                    const uint64_t b1 =
                        (uint64_t)xmax < total_bits - 1 - row ? row + (uint64_t)xmax : total_bits - 1;
                    for (uint64_t w = b0 >> 5; w <= (b1 >> 5); w++) {
                        const uint32_t lo = w == (b0 >> 5) ? (uint32_t)(b0 & 31u) : 0u;
                        const uint32_t hi = w == (b1 >> 5) ? (uint32_t)(b1 & 31u) : 31u;
                        atomicXor(table + w, (0xFFFFFFFFu >> lo) & (0xFFFFFFFFu << (31u - hi)));
                    }
                }
            }
        }
    }
    __syncthreads();  // this workgroup has read the counters
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&ws.misc[2], 1u) == gridDim.x - 1) {  // every workgroup has read them
            ws.misc[0] = 0u;
            ws.misc[1] = 0u;
            ws.misc[2] = 0u;
            __threadfence();
        }
    }
}

// Morton -> linear layout, the inverse of the brick assembly above: one
// workgroup per block of 32 rows in y x 8 rows in z (G power of two, 32..1024);
// a thread reads its 8 consecutive Morton words (two 16-byte loads), undoes the
// index swap and the nibble exchange (both are involutions), scatters the 32
// bytes into the LDS rows, and the block then stores its rows coalesced.
__global__ void __launch_bounds__(256) bricks_to_rows_kernel(const uint32_t* __restrict__ src,
                                                             uint32_t* __restrict__ dst, uint32_t G) {
    extern __shared__ uint32_t tile[];  // [8 * 32 rows][rw + 1]
    const uint32_t rw = G >> 5, nby = G >> 5, ts = rw + 1;
    const uint32_t rb = blockIdx.x;
    const uint32_t z0 = (rb / nby) * 8u, y0 = (rb % nby) * 32u;
    const uint64_t base_w = morton3(0u, y0, z0) >> 5;
    uint8_t* tb = (uint8_t*)tile;
    // all of a thread's brick loads (G / 256 <= 4 groups of 8 words) are issued first
    uint4 ld[4][2];
#pragma unroll
    for (int it = 0; it < 4; it++) {
        const uint32_t t = threadIdx.x + 256u * (uint32_t)it;
        if (t < G) {
            const uint32_t z2 = t & 1u, x3 = (t >> 1) & 1u, y3 = (t >> 2) & 1u, x4 = (t >> 3) & 1u,
                           y4 = (t >> 4) & 1u, xh = t >> 5;
            uint64_t wl = ((uint64_t)z2 << 3) | ((uint64_t)x3 << 4) | ((uint64_t)y3 << 5) | ((uint64_t)x4 << 7) |
                          ((uint64_t)y4 << 8);
            for (uint32_t b = 0; (xh >> b) != 0u; b++) wl |= (uint64_t)((xh >> b) & 1u) << (10u + 3u * b);
            const uint4* s4 = (const uint4*)(src + base_w + wl);
            ld[it][0] = s4[0];
            ld[it][1] = s4[1];
        }
    }
#pragma unroll
    for (int it = 0; it < 4; it++) {
        const uint32_t t = threadIdx.x + 256u * (uint32_t)it;
        if (t >= G) break;
        const uint32_t z2 = t & 1u, x3 = (t >> 1) & 1u, y3 = (t >> 2) & 1u, x4 = (t >> 3) & 1u,
                       y4 = (t >> 4) & 1u, xh = t >> 5;
        const uint4 lo = ld[it][0], hi = ld[it][1];
        const uint32_t out[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
        uint32_t W[8];
#pragma unroll
        for (int w = 0; w < 8; w++) {
            const uint32_t v = out[((w >> 2) & 1) | (((w >> 1) & 1) << 1) | ((w & 1) << 2)];
            const uint32_t d = ((v >> 6) ^ v) & 0x00CC00CCu;
            W[w] = v ^ d ^ (d << 6);
        }
#pragma unroll
        for (int w = 0; w < 8; w++) {
            if (w & 2) continue;
            const uint32_t d = (W[w] ^ (W[w | 2] >> 4)) & 0x0F0F0F0Fu;
            W[w] ^= d;
            W[w | 2] ^= d << 4;
        }
        const uint32_t xbyte = x3 | (x4 << 1) | (xh << 2);  // x >> 3
        const uint32_t j = xbyte >> 2, byte_in_word = 3u - (xbyte & 3u);  // MSB-first byte -> little-endian slot
#pragma unroll
        for (int q = 0; q < 32; q++) {
            const uint32_t y_l = ((uint32_t)q & 7u) | (y3 << 3) | (y4 << 4);
            const uint32_t z_l = (((uint32_t)q >> 3) & 3u) | (z2 << 2);
            const uint32_t byte = (W[q >> 2] >> (24u - 8u * ((uint32_t)q & 3u))) & 0xFFu;
            tb[4u * ((z_l * 32u + y_l) * ts + j) + byte_in_word] = (uint8_t)byte;
        }
    }
    __syncthreads();
    const uint32_t total = 256u * rw;
    if ((rw & 3u) == 0) {
        // 16-byte stores: a row is rw / 4 whole quads
        const uint32_t rq = rw >> 2, zq = 8u * rw;
        uint4* __restrict__ d4 = (uint4*)dst;
        for (uint32_t i = threadIdx.x; i < (total >> 2); i += 256) {
            const uint32_t zl = i / zq, k = i - zl * zq;
            const uint32_t* r = tile + (i / rq) * ts + (i % rq) * 4u;
            d4[((((uint64_t)(z0 + zl) * G + y0) * rw) >> 2) + k] = make_uint4(r[0], r[1], r[2], r[3]);
        }
        return;
    }
    for (uint32_t i = threadIdx.x; i < total; i += 256) {
        const uint32_t zl = i / (32u * rw), k = i % (32u * rw);
        dst[((uint64_t)(z0 + zl) * G + y0) * rw + k] = tile[(i / rw) * ts + i % rw];
    }
}

bool layout_convert_fast(uint32_t G) { return G >= 32u && G <= 1024u && (G & (G - 1)) == 0; }

// Whole-table layout conversion for a power-of-two G in 32..1024; 16-byte
// aligned tables.
hipError_t launch_layout_convert(const uint32_t* src, uint32_t* dst, uint32_t G, bool to_morton, hipStream_t stream) {
    const uint32_t nblocks = (G / 32u) * (G / 8u);
    const size_t lds = 4ull * 256u * (G / 32u + 1u);
    if (to_morton)
        marker_build_kernel<false><<<nblocks, 256, lds, stream>>>(src, dst, G, 8u, VOX_LAYOUT_MORTON, 0);
    else
        bricks_to_rows_kernel<<<nblocks, 256, lds, stream>>>(src, dst, G);
    return hipGetLastError();
}

// ---- host ----------------------------------------------------------------------
struct MarkerBuffers {
    void* ws = nullptr;       // misc + spill records
    void* scratch = nullptr;  // linear marker table (Morton output, accumulation)
    uint64_t scratch_bytes = 0;
    bool scratch_dirty = true;  // false once zero: the brick pass returns it to zero
};
static std::mutex g_mws_mutex;
static std::map<std::pair<int, hipStream_t>, MarkerBuffers> g_mws;

bool marker_supported(uint32_t G, int layout, bool accumulate) {
    (void)accumulate;
    if (G < 32 || (G & 31u) || G > 16384u) return false;  // bit indices below 2^42 (spill records)
    if (layout == VOX_LAYOUT_MORTON) return (G & (G - 1)) == 0 && G <= 4096;
    return true;
}

hipError_t launch_marker_in_place(const vox_info& info, const float* tris, uint32_t* table,
                                  unsigned long long* offrow, const MarkerWs& ws, hipStream_t stream, uint32_t big);

hipError_t launch_solid_marker(const vox_info& info, const float* tris, uint32_t* table, int layout,
                               bool accumulate, unsigned long long* offrow, hipStream_t stream, uint32_t big) {
    const uint32_t G = info.gridsize[0];
    // linear, no accumulation: markers into the scratch table and one streaming
    // pass (suffix into the table, scratch back to zero) where rows are whole
    // 16-byte groups and the scratch table can be had; in place (zero-fill,
    // markers, suffix) otherwise
    bool stream_rows = layout == VOX_LAYOUT_LINEAR && !accumulate && rows4_ok(G, table);
    bool in_place = layout == VOX_LAYOUT_LINEAR && !accumulate && !stream_rows;
    const uint64_t table_bytes = (uint64_t)G * G * G / 8;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    MarkerWs ws;
    uint32_t* markers = table;
    bool need_fill = in_place;
    {
        std::lock_guard<std::mutex> lock(g_mws_mutex);
        MarkerBuffers& b = g_mws[{dev, stream}];
        if (!b.ws) {
            e = hipMalloc(&b.ws, 256 + 16ull * kSpillCap);
            if (e == hipSuccess) e = hipMemset(b.ws, 0, 256);  // the counters start at zero
            if (e != hipSuccess) return e;
        }
        if (!in_place && b.scratch_bytes < table_bytes) {
            if (b.scratch) {
                e = hipStreamSynchronize(stream);  // a previous call on this stream may still read it
                if (e == hipSuccess) e = hipFree(b.scratch);
                b.scratch = nullptr;
                b.scratch_bytes = 0;
                if (e != hipSuccess) return e;
            }
            e = hipMalloc(&b.scratch, table_bytes);
            if (e == hipErrorOutOfMemory && stream_rows) {
                (void)hipGetLastError();  // the in-place form needs no scratch table
                b.scratch = nullptr;
                stream_rows = false;
                in_place = true;
            } else if (e != hipSuccess) {
                return e;
            } else {
                b.scratch_bytes = table_bytes;
                b.scratch_dirty = true;
            }
        }
        ws = MarkerWs{(unsigned*)b.ws, (unsigned long long*)((char*)b.ws + 256)};
        if (!in_place) {
            markers = (uint32_t*)b.scratch;
            need_fill = b.scratch_dirty;
            b.scratch_dirty = true;  // until this call's brick pass is queued
        }
    }
    if (in_place) return launch_marker_in_place(info, tris, table, offrow, ws, stream, big);
    if (need_fill) {
        e = launch_fill_zero(info, markers, stream);
        if (e != hipSuccess) return e;
    }
    const bool linear = layout == VOX_LAYOUT_LINEAR;
    const ItemMap none{nullptr, nullptr, 1, 0};
    ItemDev one{};  // markers: into the scratch table
    one.big = big;
    one.info = info;
    one.tris = tris;
    one.table = markers;
    one.offrow = offrow;
    one.layout = layout;
    one.solid = 1;
    ItemDev fin = one;  // spills: into the final table
    fin.table = table;
    if (info.n_triangles) {
        const uint64_t blocks = (info.n_triangles + kMBlock - 1) / kMBlock;
        solid_marker_kernel<false, false><<<(unsigned)blocks, kMBlock, 0, stream>>>(one, none, ws, linear ? 1 : 0,
                                                                                    RegionArgs{});
    }
    if (stream_rows) {
        const uint64_t nrows = (uint64_t)G * G;
        const uint64_t per = 64 / (G >> 7);
        uint64_t blocks = (nrows / per + 3) / 4;
        if (blocks > 262144) blocks = 262144;
        suffix_scratch_rows4_kernel<<<(unsigned)blocks, 256, 0, stream>>>(markers, table, G, nrows);
    } else {
        const uint32_t rz = linear ? 1u : (G <= 1024u ? 8u : 2u);
        const uint32_t nblocks = (G / 32u) * (G / rz);
        marker_build_kernel<true><<<nblocks, 256, 4ull * 32u * rz * (G / 32u + 1u), stream>>>(
            markers, table, G, rz, layout, accumulate ? 1 : 0);
    }
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    {
        std::lock_guard<std::mutex> lock(g_mws_mutex);
        g_mws[{dev, stream}].scratch_dirty = false;  // every row of the G^3 table was read and cleared
    }
    if (linear) spill_kernel<<<kSpillBlocks, 256, 0, stream>>>(fin, none, none, ws);
    return hipGetLastError();
}

static hipError_t marker_ws_for(hipStream_t stream, MarkerWs& ws) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lock(g_mws_mutex);
    MarkerBuffers& b = g_mws[{dev, stream}];
    if (!b.ws) {
        e = hipMalloc(&b.ws, 256 + 16ull * kSpillCap);
        if (e == hipSuccess) e = hipMemset(b.ws, 0, 256);  // the counters start at zero
        if (e != hipSuccess) return e;
    }
    ws = MarkerWs{(unsigned*)b.ws, (unsigned long long*)((char*)b.ws + 256)};
    return hipSuccess;
}

// Region algorithm, solid: markers staged per region (phase 1), the build
// pass (XOR merge, row suffix, one write per word, Morton bricks for the
// Morton layout), then the spilled runs of the linear layout.
hipError_t launch_solid_region(const vox_info& info, const float* tris, uint32_t* table, int layout,
                               bool accumulate, unsigned long long* offrow, hipStream_t stream, uint32_t big,
                               const RegionKnobs& K) {
    RegionPlan P;
    if (!make_region_plan(info.gridsize[0], layout, true, P, K)) return hipErrorInvalidValue;
    RegionWs W;
    hipError_t e = region_workspace(P, (info.n_triangles + kMBlock - 1) / kMBlock, stream, W);
    if (e != hipSuccess) return e;
    MarkerWs ws;
    e = marker_ws_for(stream, ws);
    if (e != hipSuccess) return e;
    const bool linear = layout == VOX_LAYOUT_LINEAR;
    const ItemMap none{nullptr, nullptr, 1, 0};
    ItemDev one{};
    one.big = big;
    one.info = info;
    one.tris = tris;
    one.table = table;  // spilled runs go to the final table
    one.offrow = offrow;
    one.layout = layout;
    one.solid = 1;
    if (info.n_triangles) {
        const uint64_t blocks = (info.n_triangles + kMBlock - 1) / kMBlock;
        solid_marker_kernel<false, true><<<(unsigned)blocks, kMBlock, 0, stream>>>(one, none, ws, linear ? 1 : 0,
                                                                                  RegionArgs{P, W});
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    e = launch_region_build(P, W, table, accumulate, stream);
    if (e != hipSuccess) return e;
    if (linear) spill_kernel<<<kSpillBlocks, 256, 0, stream>>>(one, none, none, ws);
    return hipGetLastError();
}

// Linear layout, no accumulation: zero-fill the table, markers into it, the
// in-place row suffix, then the spilled runs.  `ws` holds the counters (zero
// on entry; the spill pass returns them to zero) and the spill records.
hipError_t launch_marker_in_place(const vox_info& info, const float* tris, uint32_t* table,
                                  unsigned long long* offrow, const MarkerWs& ws, hipStream_t stream, uint32_t big) {
    const uint32_t G = info.gridsize[0];
    hipError_t e = launch_fill_zero(info, table, stream);
    if (e != hipSuccess) return e;
    const ItemMap none{nullptr, nullptr, 1, 0};
    ItemDev one{};
    one.big = big;
    one.info = info;
    one.tris = tris;
    one.table = table;
    one.offrow = offrow;
    one.layout = VOX_LAYOUT_LINEAR;
    one.solid = 1;
    if (info.n_triangles) {
        const uint64_t blocks = (info.n_triangles + kMBlock - 1) / kMBlock;
        solid_marker_kernel<false, false><<<(unsigned)blocks, kMBlock, 0, stream>>>(one, none, ws, 1, RegionArgs{});
    }
    if (rows4_ok(G, table)) {
        const uint64_t nrows = (uint64_t)G * G;
        const uint64_t per = 64 / (G >> 7);
        uint64_t blocks = (nrows / per + 3) / 4;
        if (blocks > 262144) blocks = 262144;
        suffix_rows4_kernel<<<(unsigned)blocks, 256, 0, stream>>>(table, G, nrows);
    } else {
        const uint64_t nrows = (uint64_t)G * G;
        const uint32_t rw = G >> 5;
        const uint64_t rows_per_wave = rw <= 64 ? 64 / rw : 1;
        const uint64_t waves = (nrows + rows_per_wave - 1) / rows_per_wave;
        uint64_t blocks = (waves + 3) / 4;
        if (blocks > 65536) blocks = 65536;
        suffix_rows_kernel<<<(unsigned)blocks, 256, 0, stream>>>(table, G, nrows);
    }
    spill_kernel<<<kSpillBlocks, 256, 0, stream>>>(one, none, none, ws);
    return hipGetLastError();
}

// ---- the allocation-free form (vox_solid_noalloc) --------------------------------
// Its scratch is this file-scope device array: the spill counters (256 bytes)
// and kSpillCap spill records, 1 MiB + 256 bytes in all, whatever the input
// (more spilled runs than records are recomputed from the triangles, which
// needs no further memory).  Each GPU holds its own copy until the process
// ends.  In host code `scratch` only names the symbol: its device address on
// the current GPU comes from hipGetSymbolAddress, taken once per device.
constexpr size_t kStaticScratchBytes = 256 + 16ull * kSpillCap;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

constexpr int kMaxStaticDevices = 64;
static std::mutex g_static_mutex;
static void* g_static_ptr[kMaxStaticDevices];  // device address of `scratch`, per device

// Zero the counter words (a kernel, so the sequence has no memset node when a
// caller captures it into a graph).
__global__ void __launch_bounds__(64) clear_counters_kernel(unsigned* misc, int words) {
    for (int i = threadIdx.x; i < words; i += 64) misc[i] = 0u;
}

size_t static_scratch_bytes() { return kStaticScratchBytes; }

hipError_t static_scratch_init(int device, void** out) {
    if (device < 0 || device >= kMaxStaticDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_static_mutex);
    if (!g_static_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));  // the current device's copy
        if (e != hipSuccess) return e;
        g_static_ptr[device] = p;
    }
    if (out) *out = g_static_ptr[device];
    return hipSuccess;
}

hipError_t launch_solid_marker_static(const vox_info& info, const float* tris, uint32_t* table,
                                      unsigned long long* offrow, hipStream_t stream) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    void* p = nullptr;
    e = static_scratch_init(dev, &p);
    if (e != hipSuccess) return e;
    const MarkerWs ws{(unsigned*)p, (unsigned long long*)((char*)p + 256)};
    clear_counters_kernel<<<1, 64, 0, stream>>>(ws.misc, 64);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_marker_in_place(info, tris, table, offrow, ws, stream, 0);
}

// ---- batched form (linear layout, no accumulation: in place) -------------------
uint64_t marker_ws_bytes() { return 256 + 16ull * kSpillCap; }
uint32_t marker_block_units(uint64_t n_triangles) { return (uint32_t)((n_triangles + kMBlock - 1) / kMBlock); }
uint32_t marker_row_units(uint32_t G) { return (uint32_t)(((uint64_t)G * G + kSuffixRows - 1) / kSuffixRows); }
bool marker_batch_supported(uint32_t G, int layout, bool accumulate) {
    return layout == VOX_LAYOUT_LINEAR && !accumulate && marker_supported(G, layout, accumulate);
}

// The tables are zero-filled beforehand (the batch's fill class).  m_blocks:
// block prefix per job, m_rows: suffix unit prefix, m_tris: triangle prefix.
// mws (marker_ws_bytes()) must be zero when first used; the spill pass
// leaves it so.
hipError_t launch_marker_jobs(const ItemMap& m_blocks, const ItemMap& m_rows, const ItemMap& m_tris,
                              uint32_t blocks, uint32_t row_units, void* mws, hipStream_t stream) {
    const MarkerWs ws{(unsigned*)mws, (unsigned long long*)((char*)mws + 256)};
    const ItemDev dummy{};
    if (blocks) solid_marker_kernel<true, false><<<blocks, kMBlock, 0, stream>>>(dummy, m_blocks, ws, 1, RegionArgs{});
    if (row_units) suffix_jobs_kernel<<<row_units, 256, 0, stream>>>(m_rows);
    spill_kernel<<<kSpillBlocks, 256, 0, stream>>>(dummy, m_blocks, m_tris, ws);
    return hipGetLastError();
}

}  // namespace voxamd
