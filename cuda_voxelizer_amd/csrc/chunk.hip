// chunk.hip -- single-pass surface voxelization with LDS word merging.
//
// One workgroup takes a chunk of 256 consecutive triangles; the fewer chunks
// share a table word, the fewer atomics, so a mesh whose consecutive
// triangles lie close together voxelizes fastest (its chunks are compact
// patches; input order is kept as given).  Each lane loads and rasterizes
// its own triangle (ZX edges once per z slice, then per row the YZ edges
// and, at the x that passed ZX, the plane and XY edges) and ORs the passing
// bits into an LDS hash table keyed by table word, so bits that several
// triangles of the chunk set in one word are merged before memory is
// touched.  Triangles with a large voxel box (or one far larger than the
// rest of their wave's) are queued; afterwards their owners publish test
// records to LDS (up to kRecs per round) and the workgroup deals the rounds'
// (row, 32-voxel x chunk) items evenly over its lanes, testing only the x
// range that the plane band leaves open.  The table is then updated with one
// non-returning global atomicOr per distinct word; words that do not fit in
// the hash go straight to global atomics.  The table is zero-filled
// beforehand (unless accumulating).
#include "jobs.h"
#include "raster.h"
#include "region.h"
#include "vox_kernels.h"

#pragma clang fp contract(off)

namespace voxamd {

constexpr int kChunk = 256;  // triangles per workgroup (one per lane)
constexpr int kHashBits = 11;
constexpr int kHashCap = 1 << kHashBits;
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
// Slots tried before a word goes straight to memory: a chunk of scattered
// triangles has more distinct words than slots, and probing a full table costs
// more than the atomic it would save.
constexpr int kProbes = 2;
constexpr int kBigVox = 256;    // voxel boxes above this (or wider than 32 in x) are queued
// ...and, with the built-in setting, boxes of more than kRowsFloor rows and
// more than kRowsRatio times the mean of their wave's boxes: one lane walking
// a box far larger than its neighbours' holds up the whole workgroup.  The
// Stanford bunny at 64^3 (5,110 triangles, 20 workgroups): the median
// workgroup's largest lane box has 40 rows against a mean of 7.4; queuing
// those 1.7 % of triangles halves it.  The icospheres, tori and displaced
// spheres of the benchmarks have no box above the floor, so their waves skip
// the rule after one ballot.
constexpr int kRowsFloor = 16;
constexpr int kRowsRatio = 3;
constexpr int kRecs = 48;       // queued triangles whose records are in LDS at once

struct ChunkShared {
    uint32_t keys[kHashCap];
    uint32_t vals[kHashCap];
    union {
        SurfRec recs[kRecs];      // records of the queued triangles of one round
        RegionFlushShared flush;  // last: grouping of the region algorithm's pairs
    };
    unsigned long long incl[kRecs];  // inclusive prefix of the round's item counts
    unsigned wq[kChunk / 64];        // queued triangles per wave
};

// ident: the slot is the key modulo kHashCap.  Used when the whole table has
// at most kHashCap words (its slots are preset to their keys, so every word is
// found at its first probe) and for Morton keys (a chunk's words are a compact
// range of brick codes); linear keys repeat modulo the table size with the z
// stride, so they are hashed multiplicatively.
// REG (region algorithm): more slots are tried, and a word that finds none is
// staged as a fallback pair instead of going to the table.
template <bool REG>
__device__ __forceinline__ void word_or(ChunkShared& S, uint32_t* __restrict__ table, bool ident, uint32_t key,
                                        uint32_t mask, const RegionArgs& RA) {
    uint32_t h = ident ? (key & (kHashCap - 1)) : (key * 2654435761u) >> (32 - kHashBits);
#pragma unroll 1
    for (int probe = 0; probe < (REG ? kRegionProbes : kProbes); probe++) {
        // one round trip: the CAS claims an empty slot or returns the key
        // already there (a read first, then a CAS on empty slots, took two for
        // every new word: 1024^3 linear torus 114 vs 108 us per call)
        const uint32_t k = atomicCAS(&S.keys[h], kEmpty, key);
        if (k == kEmpty || k == key) {
            atomicOr(&S.vals[h], mask);
            return;
        }
        h = (h + 1) & (kHashCap - 1);
    }
    if (REG)
        region_fallback(RA.P, RA.W, key, mask);
    else
        atomicOr(table + key, mask);  // hash table crowded: go to memory directly
}

// Passing voxels of row (y, z) as a mask (bit k = voxel cx + k, cx a multiple
// of 32) -> hash, keyed by table word (REG, linear layout: by region row key).
// SPEC: what is known of every job at compile time (see ChunkSpec).
template <bool REG, int SPEC>
__device__ __forceinline__ void emit_mask(ChunkShared& S, uint32_t* __restrict__ table, bool ident,
                                          const vox_info& info, int layout, int cx, int y, int z, uint32_t mask,
                                          const RegionArgs& RA) {
    if (!mask) return;
    const uint32_t G = info.gridsize[0];
    if (SPEC == kSpecMorton || (SPEC == kSpecAny && layout == VOX_LAYOUT_MORTON)) {
        // codes of grids up to 1024 fit 30 bits: 32-bit spreads
        const uint64_t base = (SPEC == kSpecMorton || G <= 1024u) ? (uint64_t)(spread10((uint32_t)cx) | (spread10((uint32_t)y) << 1) |
                                                      (spread10((uint32_t)z) << 2))
                                         : morton3((uint32_t)cx, (uint32_t)y, (uint32_t)z);
        const uint32_t yz_low = ((uint32_t)(y & 1) << 1) | ((uint32_t)(z & 1) << 2) | ((uint32_t)((y >> 1) & 1) << 4);
        morton_chunk_words(base, yz_low, mask,
                           [&](uint64_t w, uint32_t bits) { word_or<REG>(S, table, ident, (uint32_t)w, bits, RA); });
    } else if (REG) {
        word_or<REG>(S, table, ident, region_row_key(RA.P, (uint32_t)cx >> 5, (uint32_t)y, (uint32_t)z),
                     __brev(mask), RA);
    } else if (SPEC == kSpecLinear || (G & 31u) == 0) {
        // (z G + y) G/32 + cx/32 < G^3/32 < 2^32 (chunk_supported), so 32-bit arithmetic is exact
        const uint32_t w = ((uint32_t)z * G + (uint32_t)y) * (G >> 5) + ((uint32_t)cx >> 5);
        word_or<REG>(S, table, ident, w, __brev(mask), RA);
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
                if (bits) word_or<REG>(S, table, ident, cur, bits, RA);
                cur = w;
                bits = 0;
            }
            bits |= bit_mask(idx);
        }
        if (bits) word_or<REG>(S, table, ident, cur, bits, RA);
    }
}

// A lane's own box (x extent <= 32): the ZX edge tests depend on (z, x) only,
// so they are evaluated once per z slice into a mask over the box's x range;
// a row then tests the XY edges and the plane only at the x that passed ZX.
// Every test is the same binary32 expression as in vox_device.h, evaluated
// for a subset of the voxels: the resulting bits are unchanged.
template <bool REG, int SPEC>
__device__ __forceinline__ void lane_raster(ChunkShared& S, uint32_t* __restrict__ table, bool ident,
                                            const SurfTri& T, const vox_info& info, int layout,
                                            const RegionArgs& RA) {
    const float ux = info.unit[0], uy = info.unit[1], uz = info.unit[2];
    const int xa = T.lo[0], xb = T.hi[0];
    const int c0 = xa & ~31, sh = xa - c0;
    // Morton codes of grids up to 1024 fit 30 bits: the x, y and z spreads are
    // computed once per lane, row and slice instead of once per emitted mask
    const bool m32 = SPEC == kSpecMorton ||
                     (SPEC == kSpecAny && layout == VOX_LAYOUT_MORTON && info.gridsize[0] <= 1024u);
    const uint32_t sc0 = m32 ? spread10((uint32_t)c0) : 0u, sc1 = m32 ? spread10((uint32_t)c0 + 32u) : 0u;
    const auto put = [&](uint64_t w, uint32_t bits) { word_or<REG>(S, table, ident, (uint32_t)w, bits, RA); };
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
            if (m32) {
                const uint32_t syz = (spread10((uint32_t)y) << 1) | (spread10((uint32_t)z) << 2);
                const uint32_t yz_low =
                    ((uint32_t)(y & 1) << 1) | ((uint32_t)(z & 1) << 2) | ((uint32_t)((y >> 1) & 1) << 4);
                if ((uint32_t)M) morton_chunk_words(sc0 | syz, yz_low, (uint32_t)M, put);
                if (M >> 32) morton_chunk_words(sc1 | syz, yz_low, (uint32_t)(M >> 32), put);
                continue;
            }
            emit_mask<REG, SPEC>(S, table, ident, info, layout, c0, y, z, (uint32_t)M, RA);
            if (M >> 32) emit_mask<REG, SPEC>(S, table, ident, info, layout, c0 + 32, y, z, (uint32_t)(M >> 32), RA);
        }
    }
}

// Work unit: `span` consecutive chunks of kChunk triangles of one job, merged
// in one LDS table and flushed once -- or, with `parts` > 1, one chunk whose
// queued items are shared by `parts` workgroups (each does the setup, takes
// every parts-th slice of the items, and part 0 also the lane rasters; queue
// slots follow lane order, so every part sees the same items).  MULTI: the job
// comes from the launch's job map; otherwise it is `one` (kernel arguments).
// LOOP: some job has span > 1 (otherwise the chunk loop is compiled out, which
// keeps the register count of the one-chunk form down).  REG: the merged
// words are staged as region pairs (RA) instead of ORed into the table.
// Workgroups per CU: the LDS allows six.  The bound holds the registers at
// five waves per SIMD, or six with WIDE (a few spilled registers: it pays for
// the linear layout, whose chunks touch more words and so wait longer on the
// LDS table, not for the Morton layout), or four for the chunk loop.  SPEC:
// the layout when it is known at compile time, so the other layout's word
// paths are compiled out (at six waves this removes the spills).
template <bool MULTI, bool LOOP, bool WIDE, bool REG, int SPEC = kSpecAny>
__global__ void __launch_bounds__(kChunk, LOOP ? 4 : (WIDE ? 6 : 5))
    surface_chunk_kernel(ItemDev one, ItemMap M, RegionArgs RA) {
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
    const int layout = SPEC == kSpecLinear ? (int)VOX_LAYOUT_LINEAR
                       : SPEC == kSpecMorton ? (int)VOX_LAYOUT_MORTON : it.layout;
    const uint64_t n = info.n_triangles;
    const uint32_t parts = it.parts > 1u ? it.parts : 1u;
    const uint32_t span = (!LOOP || parts > 1u) ? 1u : (it.span > 1u ? it.span : 1u);
    const uint32_t part = local % parts;
    const uint64_t begin = (uint64_t)(local / parts) * span * kChunk;
    if (begin >= n) return;
    const uint64_t end = n - begin > (uint64_t)span * kChunk ? begin + (uint64_t)span * kChunk : n;
    const uint32_t G = info.gridsize[0];
    const uint64_t words = ((uint64_t)G * G * G + 31) / 32;
    const bool dense = !REG && words <= (uint64_t)kHashCap;  // region keys are not table words
    const bool ident = dense || layout == VOX_LAYOUT_MORTON;
    for (int i = threadIdx.x; i < kHashCap; i += kChunk) {
        S.keys[i] = dense && (uint64_t)i < words ? (uint32_t)i : kEmpty;
        S.vals[i] = 0u;
    }
    __syncthreads();  // every slot is set before any lane inserts into the table
    uint64_t first = begin;
    do {
        const int count = (int)(end - first < (uint64_t)kChunk ? end - first : (uint64_t)kChunk);
        // each lane loads its own triangle: the nine loads of a lane are in
        // flight together and a wave's 2304 bytes are contiguous, so they are
        // served by the cache after the first; staging the chunk through LDS
        // cost a barrier and 9 KB of LDS (512^3 Morton torus 51.2 vs 48.0 us)
        bool big = false;
        int rows = 0;  // rows of a lane-path box
        SurfTri T;
        float v[3][3];
        if ((int)threadIdx.x < count) {
            load_tri(it.tris, first + threadIdx.x, info, v);
            surface_box(v, info, T.lo, T.hi);
            const int nx = T.hi[0] - T.lo[0] + 1;
            const int ny = T.hi[1] - T.lo[1] + 1, nz = T.hi[2] - T.lo[2] + 1;
            const int64_t vol = (int64_t)nx * ny * nz;
            big = vol > (it.big ? (int64_t)it.big : (int64_t)kBigVox) || nx > 32;
            rows = big ? 0 : ny * nz;  // at most the queue threshold here (<= 2^15): no overflow
        }
        if (!it.big && __ballot(rows > kRowsFloor)) {
            // sum over the wave (dead and queued lanes add 0); the mean is
            // taken over the lanes that hold a triangle
            int sum = rows;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
            const int live = (int)__popcll(__ballot((int)threadIdx.x < count));
            if (rows > kRowsFloor && rows * live > kRowsRatio * sum) big = true;
        }
        const unsigned long long bal = __ballot(big);
        const int wave = (int)(threadIdx.x >> 6);
        if (__lane_id() == 0) S.wq[wave] = (unsigned)__popcll(bal);
        if (part == 0 && (int)threadIdx.x < count && !big) {
            surface_setup_tests(v, info, T);  // the box is T's already
            lane_raster<REG, SPEC>(S, table, ident, T, info, layout, RA);  // nx <= 32 here
        }
        __syncthreads();  // S.wq is complete
        int myq = -1, nq = 0;  // this lane's queue slot (lane order), queue length
#pragma unroll
        for (int w = 0; w < kChunk / 64; w++) {
            if (w == wave) myq = nq + (int)__popcll(bal & ((1ull << __lane_id()) - 1ull));
            nq += (int)S.wq[w];
        }
        if (!big) myq = -1;
        const float ux = info.unit[0], uy = info.unit[1], uz = info.unit[2];
        for (int r0 = 0; r0 < nq; r0 += kRecs) {
            const int nr = nq - r0 < kRecs ? nq - r0 : kRecs;
            if (myq >= r0 && myq < r0 + nr) {
                // the setup again, from memory (cheaper than keeping it in registers through the lane rasters)
                float v[3][3];
                load_tri(it.tris, first + threadIdx.x, info, v);
                SurfTri T;
                surface_setup(v, info, T);
                const int xa = T.lo[0], xb = T.hi[0], ya = T.lo[1], yb = T.hi[1], za = T.lo[2], zb = T.hi[2];
                int nk = 1, band = 0;
                float sx = 0.0f;
                band_chunks(T, info, xa, xb, yb, zb, nk, band, sx);
                publish(S.recs[myq - r0], T, xa, xb, ya, yb, za, zb, nk, band, sx);
                S.incl[myq - r0] = (xa > xb || ya > yb || za > zb)
                                       ? 0ull
                                       : (unsigned long long)(yb - ya + 1) * (unsigned long long)(zb - za + 1) *
                                             (unsigned long long)nk;
            }
            __syncthreads();
            if (threadIdx.x < 64) {  // inclusive scan of the round's counts (nr <= 64)
                const int lane = (int)threadIdx.x;
                unsigned long long c = lane < nr ? S.incl[lane] : 0ull;
#pragma unroll
                for (int d = 1; d < 64; d <<= 1) {
                    const unsigned long long o = __shfl_up(c, d);
                    if (lane >= d) c += o;
                }
                if (lane < nr) S.incl[lane] = c;
            }
            __syncthreads();
            const unsigned long long total = S.incl[nr - 1];
            for (unsigned long long j = threadIdx.x + (unsigned long long)part * kChunk; j < total;
                 j += (unsigned long long)kChunk * parts) {
                int lo = 0, hi = nr - 1;  // record: first whose inclusive prefix exceeds j
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (S.incl[mid] > j) hi = mid; else lo = mid + 1;
                }
                const SurfRec& R = S.recs[lo];
                const unsigned long long q = j - (lo ? S.incl[lo - 1] : 0ull);
                const unsigned nk = (unsigned)R.nk, ny = (unsigned)(R.yb - R.ya + 1);
                int k, y, z;
                if (q < 0x100000000ull) {  // 32-bit division whenever it suffices
                    const unsigned row = (unsigned)q / nk;
                    k = (int)((unsigned)q - row * nk);
                    y = R.ya + (int)(row % ny);
                    z = R.za + (int)(row / ny);
                } else {
                    const unsigned long long row = q / nk;
                    k = (int)(q - row * nk);
                    y = R.ya + (int)(row % ny);
                    z = R.za + (int)(row / ny);
                }
                const float py = (float)y * uy, pz = (float)z * uz;
                if (!pass3(R.yz, py, pz)) continue;
                int bx0 = R.xa, bx1 = R.xb;
                if (R.band) {
                    plane_band(R, ux, py, pz, bx0, bx1);
                    if (bx0 > bx1) continue;
                }
                const int cx = ((bx0 >> 5) + k) << 5;
                if (cx > bx1) continue;
                const int x0 = cx > bx0 ? cx : bx0, x1 = (cx + 31) < bx1 ? (cx + 31) : bx1;
                emit_mask<REG, SPEC>(S, table, ident, info, layout, cx, y, z,
                               row_mask(R.n, R.d1, R.d2, R.xy, R.zx, ux, py, pz, x0, x1, cx), RA);
            }
            __syncthreads();  // the records are rewritten by the next round
        }
        __syncthreads();  // S.wq is read above and rewritten by the next chunk
        first += kChunk;
    } while (LOOP && first < end);
    // every slot read before the first atomic (one LDS wait, not one per slot)
    static_assert(kHashCap % kChunk == 0, "whole slots per thread");
    constexpr int kPer = kHashCap / kChunk;
    uint32_t kk[kPer], vv[kPer];
#pragma unroll
    for (int u = 0; u < kPer; u++) {
        kk[u] = S.keys[threadIdx.x + u * kChunk];
        vv[u] = S.vals[threadIdx.x + u * kChunk];
    }
    if (REG) {
        region_flush<kChunk, kPer>(S.flush, kk, vv, S.keys, S.vals, RA.P, RA.W);
    } else {
#pragma unroll
        for (int u = 0; u < kPer; u++)
            if (kk[u] != kEmpty && vv[u] != 0u) atomicOr(table + kk[u], vv[u]);
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

// Chunks per workgroup.  A table of at most kHashCap words is merged densely
// in LDS, and every workgroup that touches it flushes up to all its words to
// the same few addresses: such jobs give each workgroup several chunks (while
// keeping at least 64 workgroups per job) to cut those atomics.
uint32_t chunk_span(uint32_t G, uint64_t n_triangles) {
    const uint64_t words = ((uint64_t)G * G * G + 31) / 32;
    if (words > (uint64_t)kHashCap) return 1u;
    const uint64_t chunks = (n_triangles + kChunk - 1) / kChunk;
    const uint64_t s = chunks / 64;
    return s < 1 ? 1u : (s > 16 ? 16u : (uint32_t)s);
}

// Workgroups per chunk.  A coarse mesh (fewer than G^2/4 triangles) has few
// chunks of large triangles: its queued items are split over up to 32
// workgroups per chunk, aiming at 256 workgroups per job.
uint32_t chunk_parts(uint32_t G, uint64_t n_triangles) {
    if (!coarse_mesh(n_triangles, G)) return 1u;
    const uint64_t chunks = (n_triangles + kChunk - 1) / kChunk;
    const uint64_t p = chunks ? 256 / chunks : 1;
    return p < 1 ? 1u : (p > 32 ? 32u : (uint32_t)p);
}

uint32_t chunk_units(uint64_t n_triangles, uint32_t span, uint32_t parts) {
    const uint64_t chunks = (n_triangles + kChunk - 1) / kChunk;
    if (parts > 1u) return (uint32_t)(chunks * parts);
    const uint64_t sp = span > 1u ? span : 1u;
    return (uint32_t)((chunks + sp - 1) / sp);
}

hipError_t launch_fill_jobs(const ItemDev& one, const ItemMap& M, uint32_t units, hipStream_t stream) {
    if (units == 0) return hipSuccess;
    fill_zero_kernel<<<units, 256, 0, stream>>>(one, M);
    return hipGetLastError();
}

int chunk_spec(int layout, uint32_t G) {
    if (layout == VOX_LAYOUT_LINEAR && (G & 31u) == 0) return kSpecLinear;
    if (layout == VOX_LAYOUT_MORTON && G <= 1024u) return kSpecMorton;
    return kSpecAny;
}

template <bool MULTI, bool WIDE>
static void launch_chunk_spec(int spec, uint32_t units, hipStream_t stream, const ItemDev& one, const ItemMap& M) {
    const RegionArgs none{};
    if (spec == kSpecLinear)
        surface_chunk_kernel<MULTI, false, WIDE, false, kSpecLinear><<<units, kChunk, 0, stream>>>(one, M, none);
    else if (spec == kSpecMorton)
        surface_chunk_kernel<MULTI, false, WIDE, false, kSpecMorton><<<units, kChunk, 0, stream>>>(one, M, none);
    else
        surface_chunk_kernel<MULTI, false, WIDE, false, kSpecAny><<<units, kChunk, 0, stream>>>(one, M, none);
}

// waves: 5 or 6 waves per SIMD for the one-chunk form (0: 6 for the linear
// layout, 5 for Morton -- the measured better choice for each; a layout-
// specialised form needs few enough registers to run at six either way); the
// chunk loop always runs at 4.  spec: the layout every job of M shares
// (chunk_spec of each, or kSpecAny); a single job's comes from `one`.
hipError_t launch_chunk_jobs(const ItemDev& one, const ItemMap& M, uint32_t units, bool loop, hipStream_t stream,
                             int waves, int spec) {
    if (units == 0) return hipSuccess;
    const bool wide = waves == 6 || (waves == 0 && !M.items && one.layout == VOX_LAYOUT_LINEAR);
    if (!M.items) spec = chunk_spec(one.layout, one.info.gridsize[0]);
    const RegionArgs none{};
    if (loop && M.items)
        surface_chunk_kernel<true, true, false, false><<<units, kChunk, 0, stream>>>(one, M, none);
    else if (loop)
        surface_chunk_kernel<false, true, false, false><<<units, kChunk, 0, stream>>>(one, M, none);
    else if (M.items && wide)
        launch_chunk_spec<true, true>(spec, units, stream, one, M);
    else if (M.items)
        launch_chunk_spec<true, false>(spec, units, stream, one, M);
    else if (wide)
        launch_chunk_spec<false, true>(spec, units, stream, one, M);
    else
        launch_chunk_spec<false, false>(spec, units, stream, one, M);
    return hipGetLastError();
}

static ItemDev single_job(const vox_info& info, const float* tris, uint32_t* table, int layout) {
    ItemDev one{};
    one.info = info;
    one.tris = tris;
    one.table = table;
    one.layout = layout;
    one.span = chunk_span(info.gridsize[0], info.n_triangles);
    one.parts = chunk_parts(info.gridsize[0], info.n_triangles);
    return one;
}

hipError_t launch_fill_zero(const vox_info& info, uint32_t* table, hipStream_t stream) {
    const ItemMap none{nullptr, nullptr, 1, 0};
    return launch_fill_jobs(single_job(info, nullptr, table, 0), none, fill_units(info.gridsize[0]), stream);
}

// Phase 1 of the region algorithm: one chunk of kChunk triangles per
// workgroup, its merged words staged as region pairs (no table writes).
hipError_t launch_surface_region_chunks(const vox_info& info, const float* tris, int layout, const RegionArgs& RA,
                                        hipStream_t stream, uint32_t big, int waves) {
    if (info.n_triangles == 0) return hipSuccess;
    const ItemMap none{nullptr, nullptr, 1, 0};
    ItemDev one = single_job(info, tris, nullptr, layout);
    one.span = 1;
    one.parts = 1;
    one.big = big;
    const uint32_t units = chunk_units(info.n_triangles, 1, 1);
    const bool wide = waves == 6;  // six waves per SIMD (knob); built-in: five
    if (wide)
        surface_chunk_kernel<false, false, true, true><<<units, kChunk, 0, stream>>>(one, none, RA);
    else
        surface_chunk_kernel<false, false, false, true><<<units, kChunk, 0, stream>>>(one, none, RA);
    return hipGetLastError();
}

hipError_t launch_surface_chunk(const vox_info& info, const float* tris, uint32_t* table, int layout,
                                hipStream_t stream, uint32_t big, int waves) {
    if (info.n_triangles == 0) return hipSuccess;
    const ItemMap none{nullptr, nullptr, 1, 0};
    ItemDev one = single_job(info, tris, table, layout);
    one.big = big;
    return launch_chunk_jobs(one, none, chunk_units(info.n_triangles, one.span, one.parts),
                             one.span > 1u && one.parts <= 1u, stream, waves);
}

}  // namespace voxamd
