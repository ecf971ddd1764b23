// encode.hip -- the reference's output writers (util_io.cpp) as device encoders.
//
// The writers walk the whole grid on the host and call checkVoxel once per
// voxel (O(G^3) host work, 10^9 calls at 1024^3).  Here every format is a few
// bandwidth-bound passes over the bit table:
//
//   1. permute   the linear table (a Morton table is converted first) into
//                the writer's traversal order, as a bit array:
//                  binvox      x outer, z, y fastest   (util_io.cpp:220-222)
//                  OBJ         x outer, y, z fastest   (util_io.cpp:115-117, 169-171)
//                32x32 bit blocks are transposed in registers across lanes.
//   2. count     per 32-voxel word, the output bytes (text formats) or the
//                run-length pairs (binvox) that start in the word;
//   3. scan      exclusive prefix of the counts (three-kernel device scan);
//   4. emit      every word writes its bytes at its offset.
//
// binvox runs: the reference writes (value, count) byte pairs, a new pair at
// every value change and after 255 equal voxels (util_io.cpp:218-245).  So a
// pair starts at every value change s and at s + 255k inside the run that
// starts at s.  binvox works on tiles of 2048 words (8 per thread): the
// last change position per tile, an exclusive max-scan of those (one
// workgroup) gives the run start in force at each tile, a block-level max-scan
// inside the tile the one at each thread's words; pairs are counted per tile,
// the counts scanned (one workgroup), and each tile emits its pairs from its
// offset.  Each pair's count is the distance to the next pair start (the last
// one: to G^3).
//
// Text formats follow std::ostream's default formatting (precision 6, %g):
// x + 0.5 for x < 10^5 prints as the integer digits followed by ".5".
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/cuda_voxelizer_amd/vox_capi.h"
#include "scan.h"
#include "vox_host.h"

namespace voxamd {

constexpr uint32_t kMaxEncodeGrid = 2048;  // 2^28 words; per-word offsets stay in memory

enum Order : int { ORDER_XZY = 0, ORDER_XYZ = 1 };  // slowest .. fastest axis

// ---------------------------------------------------------------------------
// 1. permutation
// ---------------------------------------------------------------------------
// 32x32 bit transpose across the 32 lanes of each half wave: lane r holds row
// r (bit 31-i = element (r, i)); afterwards lane r holds column r.  Recursive
// block swap (Hacker's Delight, transpose32) with the partner row in lane r^j.
__device__ __forceinline__ uint32_t transpose32_lanes(uint32_t a) {
    const int r = __lane_id() & 31;
    uint32_t m = 0x0000FFFFu;
#pragma unroll
    for (int j = 16; j != 0;) {
        const uint32_t o = (uint32_t)__shfl_xor((int)a, j);
        if ((r & j) == 0) {
            a ^= (a ^ (o >> j)) & m;
        } else {
            a ^= ((o ^ (a >> j)) & m) << j;
        }
        j >>= 1;
        m ^= m << j;
    }
    return a;
}

// Linear table (G % 32 == 0) -> bit array in `order`.  One workgroup per
// (fixed axis value f, group of XW words along x); its four waves walk the
// fast axis c in steps of 64 lanes.  Lane c loads the XW words of row (c, f),
// each half wave transposes 32 rows, and lane r then holds the output word of
// x = 32*w + r for 32 consecutive c.
template <int XW>
__global__ void __launch_bounds__(256) permute_kernel(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                      uint32_t G, int order) {
    const uint32_t rw = G >> 5;                     // words per x row
    const uint32_t groups = rw / XW;
    const uint32_t g = blockIdx.x % groups, f = blockIdx.x / groups;
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
    // input word of (x word xw, c) at fixed f: order XZY has f = z, c = y; XYZ has f = y, c = z
    const uint64_t base = order == ORDER_XZY ? (uint64_t)f * G * rw : (uint64_t)f * rw;
    const uint64_t cstride = order == ORDER_XZY ? (uint64_t)rw : (uint64_t)G * rw;
    for (uint32_t c0 = (uint32_t)wave * 64; c0 < G; c0 += 256) {
        const uint32_t c = c0 + (uint32_t)lane;
        uint32_t w[XW];
        if (c < G) {
            const uint32_t* src = in + base + (uint64_t)c * cstride + (uint64_t)g * XW;
            if (XW == 4) {
                const uint4 v = *(const uint4*)src;
                w[0] = v.x; w[1 % XW] = v.y; w[2 % XW] = v.z; w[3 % XW] = v.w;
            } else {
#pragma unroll
                for (int k = 0; k < XW; k++) w[k] = src[k];
            }
        } else {
#pragma unroll
            for (int k = 0; k < XW; k++) w[k] = 0u;
        }
        const uint32_t cb = (c0 >> 5) + (uint32_t)(lane >> 5);  // this half wave's block of 32 c
        const bool valid = (cb << 5) < G;
#pragma unroll
        for (int k = 0; k < XW; k++) {
            const uint32_t t = transpose32_lanes(w[k]);
            const uint64_t x = (uint64_t)(g * XW + k) * 32 + (uint64_t)(lane & 31);
            if (valid) out[(x * G + f) * rw + cb] = t;
        }
    }
}

// Any linear G: one thread per output word, 32 bit lookups.
__global__ void __launch_bounds__(256) permute_gather_kernel(const uint32_t* __restrict__ in,
                                                             uint32_t* __restrict__ out, uint32_t G, int order) {
    const uint64_t n = (uint64_t)G * G * G;
    const uint64_t words = (n + 31) / 32;
    for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < words;
         w += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t v = 0;
        for (int j = 0; j < 32; j++) {
            const uint64_t p = w * 32 + (uint64_t)j;
            if (p >= n) break;
            const uint64_t a = p / ((uint64_t)G * G), b = (p / G) % G, c = p % G;
            const uint64_t x = a, y = order == ORDER_XZY ? c : b, z = order == ORDER_XZY ? b : c;
            const uint64_t idx = x + y * G + z * (uint64_t)G * G;
            if ((in[idx >> 5] >> (31u - (uint32_t)(idx & 31u))) & 1u) v |= 1u << (31 - j);
        }
        out[w] = v;
    }
}

// ---------------------------------------------------------------------------
// 2 + 4: binvox run-length pairs
// ---------------------------------------------------------------------------
// Pair starts s + 255k (k >= 0) inside [a, b], a >= s.
__device__ __forceinline__ uint32_t pairs_in(int64_t s, int64_t a, int64_t b) {
    if (b < a) return 0u;
    const int64_t lo = (a - s + 254) / 255, hi = (b - s) / 255;
    return hi >= lo ? (uint32_t)(hi - lo + 1) : 0u;
}

// The pairs that start in one word (positions p0 .. end-1, change mask B, run
// start s in force at p0, updated to the word's last change).  Counting: their
// number, and `first` (while negative) set to the first one's start.  EMIT:
// each pair's value byte at out[2 o]; the count byte of the pair before it
// (start `prev`, offset `prev_o`, while prev >= 0) at out[2 prev_o + 1] = its
// distance to this one, and (prev, prev_o) moved to it; no byte is written at
// or past pair `cap`.  Returns their number.
template <bool EMIT>
__device__ __forceinline__ uint32_t word_pairs(uint32_t bits, uint32_t B, int64_t p0, int64_t end, int64_t& s,
                                               uint64_t o, int64_t& first, int64_t& prev, uint64_t& prev_o,
                                               uint64_t cap, uint8_t* __restrict__ out) {
    int64_t a = p0;
    uint32_t count = 0;
    for (;;) {
        const int i = B ? __clz(B) : 32;  // next change offset in the word
        const int64_t b = (i < 32 ? p0 + i : end) - 1;
        if (b >= a) {
            if (EMIT) {
                for (int64_t q = s + ((a - s + 254) / 255) * 255; q <= b; q += 255) {
                    if (o < cap) out[2 * o] = (uint8_t)((bits >> (31 - (int)(q - p0))) & 1u);
                    if (prev >= 0 && prev_o < cap) out[2 * prev_o + 1] = (uint8_t)(q - prev);
                    prev = q;
                    prev_o = o;
                    o++;
                    count++;
                }
            } else {
                const uint32_t c = pairs_in(s, a, b);
                if (c && first < 0) first = s + ((a - s + 254) / 255) * 255;
                count += c;
            }
        }
        if (i >= 32) break;
        B &= ~(0x80000000u >> i);
        s = p0 + i;
        a = s;
    }
    return count;
}

// Tiles of kRleTile words, kRlePer consecutive words per thread: a thread's
// words, their change masks and its last change position.
constexpr int kRlePer = 8;
constexpr uint64_t kRleTile = 256ull * kRlePer;

// Change mask B of word w: bit 31-i set iff voxel 32w+i differs from the
// voxel before it (voxel 0 always starts a run); positions >= n cut off.
struct RleWords {
    uint32_t bits[kRlePer], B[kRlePer];
    int64_t last;  // last change position in the thread's words, or -1
};

__device__ __forceinline__ void rle_words(const uint32_t* __restrict__ P, uint64_t words, uint64_t n, uint64_t w0,
                                          RleWords& R) {
    const uint32_t before = (w0 > 0 && w0 - 1 < words) ? P[w0 - 1] : 0u;
#pragma unroll
    for (int u = 0; u < kRlePer; u++) R.bits[u] = w0 + u < words ? P[w0 + u] : 0u;  // all loads first
    R.last = -1;
#pragma unroll
    for (int u = 0; u < kRlePer; u++) {
        const uint64_t w = w0 + u;
        const uint32_t bits = R.bits[u];
        const uint32_t prev = u ? (R.bits[u - 1] & 1u) : (w0 ? (before & 1u) : ((~bits >> 31) & 1u));
        uint32_t B = bits ^ ((bits >> 1) | (prev << 31));
        if (w >= words) {
            B = 0u;
        } else {
            const uint64_t left = n - w * 32;
            if (left < 32) B &= ~(0xFFFFFFFFu >> left);  // keep the first `left` positions
        }
        R.B[u] = B;
        if (B) R.last = (int64_t)(w * 32 + 32 - (uint64_t)__ffs(B));
    }
}

// Pass 1: last change position of each tile.
__global__ void __launch_bounds__(256) rle_tile_last_kernel(const uint32_t* __restrict__ P, uint64_t words, uint64_t n,
                                                            int64_t* __restrict__ tile_last) {
    RleWords R;
    rle_words(P, words, n, blockIdx.x * kRleTile + (uint64_t)threadIdx.x * kRlePer, R);
    uint64_t total;
    block_excl_scan<MaxOp>((uint64_t)R.last, &total);
    if (threadIdx.x == 0) tile_last[blockIdx.x] = (int64_t)total;
}

// Pass 3: with the run start in force at each tile (exclusive max-scan of
// pass 1), the pairs of each tile; each thread keeps the run start in force at
// its first word, its pair offset inside the tile and the start of its first
// pair for pass 5.  (Pairs start at most 255 positions apart, so every thread's
// 256 positions hold one, except past the end of the table.)
__global__ void __launch_bounds__(256) rle_tile_count_kernel(const uint32_t* __restrict__ P, uint64_t words, uint64_t n,
                                                             const int64_t* __restrict__ tile_s_in,
                                                             uint64_t* __restrict__ tile_cnt,
                                                             int64_t* __restrict__ thread_s,
                                                             uint32_t* __restrict__ thread_off,
                                                             int64_t* __restrict__ thread_first) {
    const uint64_t w0 = blockIdx.x * kRleTile + (uint64_t)threadIdx.x * kRlePer;
    RleWords R;
    rle_words(P, words, n, w0, R);
    uint64_t total;
    const int64_t ex = (int64_t)block_excl_scan<MaxOp>((uint64_t)R.last, &total);
    const int64_t s_in = tile_s_in[blockIdx.x] > ex ? tile_s_in[blockIdx.x] : ex;
    int64_t s = s_in, first = -1, prev = -1;
    uint64_t prev_o = 0;
    uint32_t cnt = 0;
#pragma unroll
    for (int u = 0; u < kRlePer; u++) {
        const uint64_t w = w0 + u;
        if (w >= words) break;
        const int64_t end = (int64_t)((n - w * 32) < 32 ? n : w * 32 + 32);
        cnt += word_pairs<false>(R.bits[u], R.B[u], (int64_t)(w * 32), end, s, 0, first, prev, prev_o, 0, nullptr);
    }
    uint64_t tile_total;
    const uint64_t o0 = block_excl_scan<SumOp>(cnt, &tile_total);
    const uint64_t g = blockIdx.x * 256ull + threadIdx.x;
    thread_s[g] = s_in;
    thread_off[g] = (uint32_t)o0;
    thread_first[g] = first;
    if (threadIdx.x == 0) tile_cnt[blockIdx.x] = tile_total;
}

// Pass 5: each thread emits the (value, count) pairs of its words from its
// tile's offset (exclusive sum-scan of the counts) plus its offset inside the
// tile; its last pair runs to the next thread's first pair (or to the end of
// the table).  Pairs past `cap` are not written (the caller re-runs the pass
// with a larger buffer).
__global__ void __launch_bounds__(256) rle_tile_emit_kernel(const uint32_t* __restrict__ P, uint64_t words, uint64_t n,
                                                            const int64_t* __restrict__ thread_s,
                                                            const uint32_t* __restrict__ thread_off,
                                                            const int64_t* __restrict__ thread_first,
                                                            uint64_t nthreads, const uint64_t* __restrict__ tile_off,
                                                            uint64_t cap, uint8_t* __restrict__ out) {
    const uint64_t w0 = blockIdx.x * kRleTile + (uint64_t)threadIdx.x * kRlePer;
    if (w0 >= words) return;
    const uint64_t g = blockIdx.x * 256ull + threadIdx.x;
    const uint64_t o0 = tile_off[blockIdx.x] + thread_off[g];
    const int64_t nf = g + 1 < nthreads ? thread_first[g + 1] : -1;
    if (o0 >= cap) return;
    RleWords R;
    rle_words(P, words, n, w0, R);
    int64_t s = thread_s[g], first = -1, prev = -1;
    uint64_t o = o0, prev_o = 0;
#pragma unroll
    for (int u = 0; u < kRlePer; u++) {
        const uint64_t w = w0 + u;
        if (w >= words) break;
        const int64_t end = (int64_t)((n - w * 32) < 32 ? n : w * 32 + 32);
        o += word_pairs<true>(R.bits[u], R.B[u], (int64_t)(w * 32), end, s, o, first, prev, prev_o, cap, out);
    }
    if (prev >= 0 && prev_o < cap) out[2 * prev_o + 1] = (uint8_t)((nf >= 0 ? nf : (int64_t)n) - prev);
}

// ---------------------------------------------------------------------------
// 2 + 4: OBJ text
// ---------------------------------------------------------------------------
__device__ __forceinline__ int ndigits(uint32_t v) {
    int d = 1;
    while (v >= 10) {
        v /= 10;
        d++;
    }
    return d;
}

__device__ __forceinline__ uint8_t* put_uint(uint8_t* o, uint32_t v) {
    const int d = ndigits(v);
    for (int k = d - 1; k >= 0; k--) {
        o[k] = (uint8_t)('0' + v % 10);
        v /= 10;
    }
    return o + d;
}

// the 12 faces of write_cube (util_io.cpp:45-90), relative vertex indices
__constant__ char kCubeFaces[] =
    "f -1 -3 -4\nf -1 -4 -2\nf -4 -3 -6\nf -4 -6 -5\nf -3 -1 -8\nf -3 -8 -6\n"
    "f -1 -2 -7\nf -1 -7 -8\nf -2 -4 -5\nf -2 -5 -7\nf -5 -6 -8\nf -5 -8 -7\n";
constexpr int kCubeFaceBytes = 12 * 11;

// vertex order of write_cube: v8, v7, ..., v1 as (dx, dy, dz) offsets
__constant__ uint8_t kCubeVerts[8][3] = {{1, 1, 0}, {0, 1, 0}, {1, 0, 0}, {0, 0, 0},
                                         {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};

__device__ __forceinline__ uint32_t obj_bytes(int fmt, uint32_t x, uint32_t y, uint32_t z) {
    if (fmt == VOX_ENC_OBJ_POINTS) return (uint32_t)(ndigits(x) + ndigits(y) + ndigits(z) + 11);
    uint32_t b = kCubeFaceBytes;
    for (int v = 0; v < 8; v++)
        b += (uint32_t)(ndigits(x + kCubeVerts[v][0]) + ndigits(y + kCubeVerts[v][1]) +
                        ndigits(z + kCubeVerts[v][2]) + 5);
    return b;
}

__device__ __forceinline__ uint8_t* obj_write(int fmt, uint8_t* o, uint32_t x, uint32_t y, uint32_t z) {
    if (fmt == VOX_ENC_OBJ_POINTS) {
        *o++ = 'v';
        const uint32_t c[3] = {x, y, z};
        for (int k = 0; k < 3; k++) {
            *o++ = ' ';
            o = put_uint(o, c[k]);
            *o++ = '.';
            *o++ = '5';
        }
        *o++ = '\n';
        return o;
    }
    for (int v = 0; v < 8; v++) {
        *o++ = 'v';
        *o++ = ' ';
        o = put_uint(o, x + kCubeVerts[v][0]);
        *o++ = ' ';
        o = put_uint(o, y + kCubeVerts[v][1]);
        *o++ = ' ';
        o = put_uint(o, z + kCubeVerts[v][2]);
        *o++ = '\n';
    }
    for (int k = 0; k < kCubeFaceBytes; k++) *o++ = (uint8_t)kCubeFaces[k];
    return o;
}

// P in XYZ order (z fastest).  EMIT = false: bytes per word -> cnt.
template <bool EMIT>
__global__ void __launch_bounds__(256) obj_kernel(const uint32_t* __restrict__ P, uint64_t words, uint32_t G, int fmt,
                                                  uint32_t* __restrict__ cnt, const uint64_t* __restrict__ off,
                                                  uint8_t* __restrict__ out) {
    const uint64_t GG = (uint64_t)G * G;
    for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < words;
         w += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t bits = P[w];
        uint32_t bytes = 0;
        uint8_t* o = EMIT ? out + off[w] : nullptr;
        while (bits) {
            const int i = __clz(bits);
            bits &= ~(0x80000000u >> i);
            const uint64_t p = w * 32 + (uint64_t)i;
            const uint32_t x = (uint32_t)(p / GG), y = (uint32_t)((p / G) % G), z = (uint32_t)(p % G);
            if (EMIT)
                o = obj_write(fmt, o, x, y, z);
            else
                bytes += obj_bytes(fmt, x, y, z);
        }
        if (!EMIT) cnt[w] = bytes;
    }
}

// ---------------------------------------------------------------------------
// 2 + 4: MagicaVoxel .vox (write_vox, util_io.cpp:249-285, through the
// vendored VoxWriter, libs/magicavoxel_file_writer/VoxWriter.cpp)
// ---------------------------------------------------------------------------
// The writer is fed AddVoxel(x, G - z, y) in x, y, z order (z fastest) and
// splits the volume into cubes of 126^3: cube (X/126, Y/126, Z/126), each
// voxel stored as the bytes (X%126, Y%126, Z%126, colour 1) in the cube's
// XYZI chunk, in feeding order; cubes are written in order of their first
// voxel.  On the device a segment is one (x, y) row of one cube: the z range
// whose Y = G - z falls in the cube.  Segments are numbered cube-major,
//   i = (c * 126 + x % 126) * 126 + y % 126,  c = (X/126 * ny + Y/126) * nz + Z/126,
// so an exclusive scan of the per-segment counts gives every voxel's index in
// its cube after subtracting the scan value at the cube's first segment.
constexpr uint32_t kVoxEdge = 126;
constexpr uint64_t kVoxSegs = (uint64_t)kVoxEdge * kVoxEdge;  // segments per cube

struct VoxGeom {
    uint32_t G, nx, ny, nz;
    uint64_t nseg;
};

__device__ __forceinline__ bool vox_segment(const VoxGeom& g, uint64_t i, uint32_t& c, uint32_t& x, uint32_t& y,
                                            uint32_t& zlo, uint32_t& zhi) {
    c = (uint32_t)(i / kVoxSegs);
    const uint32_t r = (uint32_t)(i % kVoxSegs);
    const uint32_t oz = c % g.nz, oy = (c / g.nz) % g.ny, ox = c / (g.nz * g.ny);
    x = ox * kVoxEdge + r / kVoxEdge;
    y = oz * kVoxEdge + r % kVoxEdge;
    if (x >= g.G || y >= g.G) return false;
    // Y = G - z in [126 oy, 126 oy + 125] and in [1, G]
    const uint32_t ymax = oy * kVoxEdge + kVoxEdge - 1 < g.G ? oy * kVoxEdge + kVoxEdge - 1 : g.G;
    const uint32_t ymin = oy * kVoxEdge > 1u ? oy * kVoxEdge : 1u;
    if (ymin > ymax) return false;
    zlo = g.G - ymax;
    zhi = g.G - ymin;
    return true;
}

// Bits [a, b] of P: count, first and last set position (when count > 0).
__device__ __forceinline__ uint32_t range_bits(const uint32_t* P, uint64_t a, uint64_t b, uint64_t& first,
                                               uint64_t& last) {
    uint32_t n = 0;
    first = ~0ull;
    last = 0;
    for (uint64_t w = a >> 5; w <= (b >> 5); w++) {
        uint32_t m = P[w];
        const uint64_t w0 = w << 5;
        if (w0 < a) m &= 0xFFFFFFFFu >> (uint32_t)(a - w0);
        if (w0 + 31 > b) m &= ~(0xFFFFFFFFu >> (uint32_t)(b - w0 + 1));
        if (m) {
            if (first == ~0ull) first = w0 + (uint64_t)__clz(m);
            last = w0 + 31 - (uint64_t)(__ffs(m) - 1);
            n += (uint32_t)__popc(m);
        }
    }
    return n;
}

// misc[0..5]: min x, max x, min y, max y, min z, max z; misc[6]: last position
__global__ void __launch_bounds__(256) vox_count_kernel(const uint32_t* __restrict__ P, VoxGeom g,
                                                        uint32_t* __restrict__ cnt,
                                                        unsigned long long* __restrict__ first,
                                                        unsigned* __restrict__ total,
                                                        unsigned long long* __restrict__ misc) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < g.nseg;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t c, x, y, zlo, zhi, n = 0;
        if (vox_segment(g, i, c, x, y, zlo, zhi)) {
            const uint64_t base = ((uint64_t)x * g.G + y) * g.G;
            uint64_t pf, pl;
            n = range_bits(P, base + zlo, base + zhi, pf, pl);
            if (n) {
                atomicMin(first + c, (unsigned long long)pf);
                atomicAdd(total + c, n);
                atomicMin(misc + 0, (unsigned long long)x);
                atomicMax(misc + 1, (unsigned long long)x);
                atomicMin(misc + 2, (unsigned long long)y);
                atomicMax(misc + 3, (unsigned long long)y);
                atomicMin(misc + 4, (unsigned long long)(pf - base));
                atomicMax(misc + 5, (unsigned long long)(pl - base));
                atomicMax(misc + 6, (unsigned long long)pl);
            }
        }
        cnt[i] = n;
    }
}

__global__ void __launch_bounds__(256) vox_init_kernel(unsigned long long* first, unsigned* total, uint32_t ncubes,
                                                       unsigned long long* misc) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ncubes) {
        first[i] = ~0ull;
        total[i] = 0u;
    }
    if (i < 7) misc[i] = (i & 1u) || i == 6 ? 0ull : ~0ull;
}

// voxel bytes of every segment at base[c] + 4 * off[i]
__global__ void __launch_bounds__(256) vox_emit_kernel(const uint32_t* __restrict__ P, VoxGeom g,
                                                       const uint32_t* __restrict__ cnt,
                                                       const uint64_t* __restrict__ off,
                                                       const int64_t* __restrict__ base, uint8_t* __restrict__ out) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < g.nseg;
         i += (uint64_t)gridDim.x * blockDim.x) {
        if (!cnt[i]) continue;
        uint32_t c, x, y, zlo, zhi;
        if (!vox_segment(g, i, c, x, y, zlo, zhi)) continue;
        uint8_t* o = out + (base[c] + 4 * (int64_t)off[i]);
        const uint64_t row = ((uint64_t)x * g.G + y) * g.G;
        const uint8_t bx = (uint8_t)(x % kVoxEdge), bz = (uint8_t)(y % kVoxEdge);
        for (uint32_t z = zlo; z <= zhi; z++) {
            const uint64_t p = row + z;
            if ((P[p >> 5] >> (31u - (uint32_t)(p & 31u))) & 1u) {
                o[0] = bx;
                o[1] = (uint8_t)((g.G - z) % kVoxEdge);
                o[2] = bz;
                o[3] = 1;
                o += 4;
            }
        }
    }
}

// SIZE chunk + XYZI chunk header of each written cube, before its voxels
__global__ void __launch_bounds__(256) vox_chunk_kernel(const int64_t* __restrict__ data_off,
                                                        const unsigned* __restrict__ total, uint32_t ncubes,
                                                        uint8_t* __restrict__ out) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncubes || data_off[c] < 0) return;
    const int32_t n = (int32_t)total[c];
    const int32_t hdr[10] = {0x455A4953 /* SIZE */, 12, 0, (int32_t)kVoxEdge, (int32_t)kVoxEdge, (int32_t)kVoxEdge,
                             0x495A5958 /* XYZI */, 4 * (1 + n), 0, n};
    uint8_t* o = out + data_off[c] - 40;
    for (int k = 0; k < 40; k++) o[k] = (uint8_t)(hdr[k >> 2] >> (8 * (k & 3)));
}

static unsigned grid_for(uint64_t n) {
    uint64_t b = (n + 255) / 256;
    if (b < 1) b = 1;
    if (b > 16384) b = 16384;
    return (unsigned)b;
}

}  // namespace voxamd

using namespace voxamd;

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
struct vox_encoder {
    hipStream_t stream = nullptr;
    int device = 0;
    // grow-only device buffers
    void* lin = nullptr;      uint64_t lin_bytes = 0;     // Morton -> linear copy
    void* perm = nullptr;     uint64_t perm_bytes = 0;    // permuted bits
    void* cnt = nullptr;      uint64_t cnt_bytes = 0;     // uint32 per word
    void* off = nullptr;      uint64_t off_bytes = 0;     // uint64 per word
    void* aux = nullptr;      uint64_t aux_bytes = 0;     // int64 per word (run starts)
    void* sums = nullptr;     uint64_t sums_bytes = 0;    // scan block sums
    void* out = nullptr;      uint64_t out_bytes = 0;     // encoded bytes
    uint64_t* host = nullptr;                             // pinned: 2 words read back
    uint64_t result = 0;
};

static vox_status grow(void*& p, uint64_t& have, uint64_t need) {
    if (need <= have) return VOX_OK;
    if (p) {
        hipError_t e = hipFree(p);
        p = nullptr;
        have = 0;
        if (e != hipSuccess) return hip_fail(e, "hipFree");
    }
    hipError_t e = hipMalloc(&p, need);
    if (e != hipSuccess) {
        p = nullptr;
        return fail(VOX_ERR_OOM, "encoder buffer of " + std::to_string(need) + " bytes: " + hipGetErrorString(e));
    }
    have = need;
    return VOX_OK;
}

#define ENC_CHECK(expr)                         \
    do {                                        \
        vox_status s_ = (expr);                 \
        if (s_ != VOX_OK) return s_;            \
    } while (0)
#define ENC_HIP(expr, what)                     \
    do {                                        \
        hipError_t e_ = (expr);                 \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

// ---- .vox file assembly (host) ----------------------------------------------
// The scene graph and palette that VoxWriter::SaveToFile writes after the
// cubes (VoxWriter.cpp:459-560), restated with its integer and float types:
// cube offsets use size_t differences, a float (cube - min + 0.5f) * 126,
// then double subtractions of the voxel bounding box; note that the writer
// updates minCubeY and minCubeZ as min(minCubeX, oy / oz of the latest
// voxel) (VoxWriter.cpp:450-452), so they depend on the last voxel fed.
// Float-to-int conversions outside the int range give INT_MIN, as the x86
// truncating conversion does.  The writer leaves palette entry 255
// uninitialised (VoxWriter.cpp:536-545); it is written as 0 here.
static int32_t to_int_x86(double d) {
    if (!(d >= -2147483648.0 && d < 2147483648.0)) return INT32_MIN;
    return (int32_t)d;
}

struct Blob {
    std::vector<uint8_t> b;
    void i32(int32_t v) {
        for (int k = 0; k < 4; k++) b.push_back((uint8_t)((uint32_t)v >> (8 * k)));
    }
    void str(const std::string& s) {
        i32((int32_t)s.size());
        b.insert(b.end(), s.begin(), s.end());
    }
};

static int32_t dict_size(const std::vector<std::pair<std::string, std::string>>& d) {
    int32_t n = 4;
    for (auto& kv : d) n += 8 + (int32_t)kv.first.size() + (int32_t)kv.second.size();
    return n;
}

static void dict(Blob& o, const std::vector<std::pair<std::string, std::string>>& d) {
    o.i32((int32_t)d.size());
    for (auto& kv : d) {
        o.str(kv.first);
        o.str(kv.second);
    }
}

static vox_status encode_vox(vox_encoder* e, const uint32_t* P, uint32_t G) {
    hipStream_t st = e->stream;
    VoxGeom g;
    g.G = G;
    g.nx = (G + kVoxEdge - 1) / kVoxEdge;  // X = x <= G - 1
    g.ny = G / kVoxEdge + 1;               // Y = G - z in [1, G]
    g.nz = g.nx;                           // Z = y
    const uint32_t ncubes = g.nx * g.ny * g.nz;
    g.nseg = (uint64_t)ncubes * kVoxSegs;
    const uint64_t nb = (g.nseg + kScanTile - 1) / kScanTile;
    ENC_CHECK(grow(e->cnt, e->cnt_bytes, g.nseg * 4));
    ENC_CHECK(grow(e->off, e->off_bytes, g.nseg * 8));
    ENC_CHECK(grow(e->sums, e->sums_bytes, (nb + 1) * 8));
    // aux: first (u64) | data_off (i64) | base (i64) | misc (7 u64) | total (u32)
    ENC_CHECK(grow(e->aux, e->aux_bytes, (uint64_t)ncubes * 28 + 64));
    unsigned long long* first = (unsigned long long*)e->aux;
    int64_t* data_off = (int64_t*)(first + ncubes);
    int64_t* base = data_off + ncubes;
    unsigned long long* misc = (unsigned long long*)(base + ncubes);
    unsigned* total = (unsigned*)(misc + 8);
    uint32_t* cnt = (uint32_t*)e->cnt;
    uint64_t* off = (uint64_t*)e->off;

    vox_init_kernel<<<grid_for(ncubes > 8 ? ncubes : 8), 256, 0, st>>>(first, total, ncubes, misc);
    vox_count_kernel<<<grid_for(g.nseg), 256, 0, st>>>(P, g, cnt, first, total, misc);
    ENC_HIP((device_scan<SumOp, uint32_t>(cnt, g.nseg, (uint64_t*)e->sums, off, st)), "sum-scan");
    std::vector<unsigned long long> h_first(ncubes), h_misc(7);
    std::vector<unsigned> h_total(ncubes);
    ENC_HIP(hipMemcpyAsync(h_first.data(), first, 8ull * ncubes, hipMemcpyDeviceToHost, st), "read-back");
    ENC_HIP(hipMemcpyAsync(h_total.data(), total, 4ull * ncubes, hipMemcpyDeviceToHost, st), "read-back");
    ENC_HIP(hipMemcpyAsync(h_misc.data(), misc, 8ull * 7, hipMemcpyDeviceToHost, st), "read-back");
    ENC_HIP(hipStreamSynchronize(st), "hipStreamSynchronize");

    // cubes in order of their first voxel
    std::vector<uint32_t> order;
    for (uint32_t c = 0; c < ncubes; c++)
        if (h_total[c]) order.push_back(c);
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return h_first[a] < h_first[b]; });
    std::vector<uint64_t> id_off(ncubes, 0);  // voxels of cubes with a smaller id (scan order)
    for (uint32_t c = 1; c < ncubes; c++) id_off[c] = id_off[c - 1] + h_total[c - 1];
    std::vector<int64_t> h_data(ncubes, -1), h_base(ncubes, 0);
    int64_t pos = 20;  // "VOX " 150 "MAIN" 0 childSize
    for (uint32_t c : order) {
        pos += 40;
        h_data[c] = pos;
        h_base[c] = pos - 4 * (int64_t)id_off[c];
        pos += 4 * (int64_t)h_total[c];
    }

    // scene graph + palette
    Blob tail;
    const int32_t count = (int32_t)order.size();
    std::vector<std::pair<std::string, std::string>> none;
    // root nTRN
    tail.i32(0x4E52546E);  // nTRN
    tail.i32(20 + dict_size(none) + dict_size(none));
    tail.i32(0);
    tail.i32(0); dict(tail, none); tail.i32(1); tail.i32(-1); tail.i32(-1); tail.i32(1); dict(tail, none);
    // root nGRP
    tail.i32(0x5052476E);  // nGRP
    tail.i32(4 * (2 + count) + dict_size(none));
    tail.i32(0);
    tail.i32(1); dict(tail, none); tail.i32(count);
    for (int32_t k = 0; k < count; k++) tail.i32(2 + 2 * k);
    if (count) {
        const uint64_t pl = h_misc[6];
        const uint64_t ly = (pl / G) % G, lz = pl % G;  // last voxel fed (its y and z)
        size_t minx = ~(size_t)0;
        for (uint32_t c : order) {
            const size_t ox = c / (g.nz * g.ny);
            minx = ox < minx ? ox : minx;
        }
        const size_t oy_last = (size_t)((G - lz) / kVoxEdge), oz_last = (size_t)(ly / kVoxEdge);
        const size_t miny = minx < oy_last ? minx : oy_last, minz = minx < oz_last ? minx : oz_last;
        // voxel-coordinate bounds: X = x, Y = G - z, Z = y
        const double lox = (double)h_misc[0], hix = (double)h_misc[1];
        const double loy = (double)(G - h_misc[5]), hiy = (double)(G - h_misc[4]);
        const double sx = hix - lox, sy = hiy - loy;
        for (int32_t k = 0; k < count; k++) {
            const uint32_t c = order[k];
            const size_t oz = c % g.nz, oy = (c / g.nz) % g.ny, ox = c / (g.nz * g.ny);
            const float fx = ((float)(size_t)(ox - minx) + 0.5f) * (float)(size_t)kVoxEdge;
            const float fy = ((float)(size_t)(oy - miny) + 0.5f) * (float)(size_t)kVoxEdge;
            const float fz = ((float)(size_t)(oz - minz) + 0.5f) * (float)(size_t)kVoxEdge;
            const int32_t tx = to_int_x86(std::floor((double)fx - lox - sx * 0.5));
            const int32_t ty = to_int_x86(std::floor((double)fy - loy - sy * 0.5));
            const int32_t tz = to_int_x86((double)std::floor(fz));
            std::vector<std::pair<std::string, std::string>> frame{
                {"_t", std::to_string(tx) + " " + std::to_string(ty) + " " + std::to_string(tz)}};
            tail.i32(0x4E52546E);  // nTRN
            tail.i32(20 + dict_size(none) + dict_size(frame));
            tail.i32(0);
            tail.i32(2 + 2 * k); dict(tail, none); tail.i32(3 + 2 * k); tail.i32(-1); tail.i32(0); tail.i32(1);
            dict(tail, frame);
            std::vector<std::pair<std::string, std::string>> model{{"_f", "0"}};
            tail.i32(0x5048536E);  // nSHP
            tail.i32(8 + dict_size(none) + 4 + dict_size(model));
            tail.i32(0);
            tail.i32(3 + 2 * k); dict(tail, none); tail.i32(1); tail.i32(k); dict(tail, model);
        }
    }
    tail.i32(0x41424752);  // RGBA
    tail.i32(1024);
    tail.i32(0);
    tail.i32(0x00FFFFFF);  // AddColor(255, 255, 255, 0, 0)
    for (int k = 1; k < 256; k++) tail.i32(0);
    const int64_t file_bytes = pos + (int64_t)tail.b.size();
    Blob head;
    head.i32(0x20584F56);  // "VOX "
    head.i32(150);
    head.i32(0x4E49414D);  // MAIN
    head.i32(0);
    head.i32((int32_t)(file_bytes - 20));

    ENC_CHECK(grow(e->out, e->out_bytes, (uint64_t)file_bytes + 16));
    uint8_t* out = (uint8_t*)e->out;
    ENC_HIP(hipMemcpyAsync(data_off, h_data.data(), 8ull * ncubes, hipMemcpyHostToDevice, st), "upload");
    ENC_HIP(hipMemcpyAsync(base, h_base.data(), 8ull * ncubes, hipMemcpyHostToDevice, st), "upload");
    ENC_HIP(hipMemcpyAsync(out, head.b.data(), head.b.size(), hipMemcpyHostToDevice, st), "upload");
    ENC_HIP(hipMemcpyAsync(out + pos, tail.b.data(), tail.b.size(), hipMemcpyHostToDevice, st), "upload");
    vox_emit_kernel<<<grid_for(g.nseg), 256, 0, st>>>(P, g, cnt, off, base, out);
    vox_chunk_kernel<<<(ncubes + 255) / 256, 256, 0, st>>>(data_off, total, ncubes, out);
    ENC_HIP(hipGetLastError(), "vox kernels");
    // the host vectors above are copy sources: wait before they go out of scope
    ENC_HIP(hipStreamSynchronize(st), "hipStreamSynchronize");
    e->result = (uint64_t)file_bytes;
    return VOX_OK;
}

extern "C" {

vox_status vox_encoder_create(void* stream, vox_encoder** out) {
    if (!out) return fail(VOX_ERR_INVALID, "NULL output handle");
    *out = nullptr;
    vox_encoder* e = new vox_encoder();
    e->stream = (hipStream_t)stream;
    hipError_t he = hipGetDevice(&e->device);
    if (he == hipSuccess) he = hipHostMalloc((void**)&e->host, 2 * sizeof(uint64_t), hipHostMallocDefault);
    if (he != hipSuccess) {
        delete e;
        return hip_fail(he, "encoder set-up");
    }
    *out = e;
    return VOX_OK;
}

vox_status vox_encoder_destroy(vox_encoder* e) {
    if (!e) return VOX_OK;
    hipError_t he = hipStreamSynchronize(e->stream);
    void* bufs[] = {e->lin, e->perm, e->cnt, e->off, e->aux, e->sums, e->out};
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    if (e->host) (void)hipHostFree(e->host);
    delete e;
    return he == hipSuccess ? VOX_OK : hip_fail(he, "hipStreamSynchronize");
}

vox_status vox_encode(vox_encoder* e, const uint32_t* d_table, uint32_t grid, int layout, int format,
                      uint64_t* out_bytes) {
    if (!e || !d_table || !out_bytes) return fail(VOX_ERR_INVALID, "NULL argument");
    if (format != VOX_ENC_BINVOX && format != VOX_ENC_OBJ_POINTS && format != VOX_ENC_OBJ_CUBES &&
        format != VOX_ENC_VOX)
        return fail(VOX_ERR_INVALID, "unknown encoder format");
    if (grid == 0 || grid > kMaxEncodeGrid)
        return fail(VOX_ERR_UNSUPPORTED, "encoders support grid edges 1..2048");
    if (layout != VOX_LAYOUT_LINEAR && layout != VOX_LAYOUT_MORTON)
        return fail(VOX_ERR_INVALID, "layout must be VOX_LAYOUT_LINEAR or VOX_LAYOUT_MORTON");
    if (layout == VOX_LAYOUT_MORTON && (grid < 4 || (grid & (grid - 1)) != 0))
        return fail(VOX_ERR_UNSUPPORTED, "Morton tables need a power-of-two grid >= 4");
    hipStream_t st = e->stream;
    const uint32_t G = grid;
    const uint64_t n = (uint64_t)G * G * G;
    const uint64_t words = (n + 31) / 32;
    const int order = format == VOX_ENC_BINVOX ? ORDER_XZY : ORDER_XYZ;  // .vox is fed in x, y, z order

    const uint32_t* lin = d_table;
    if (layout == VOX_LAYOUT_MORTON) {
        ENC_CHECK(grow(e->lin, e->lin_bytes, words * 4));
        ENC_CHECK(vox_convert_layout(d_table, (uint32_t*)e->lin, G, 0, st));
        lin = (const uint32_t*)e->lin;
    }
    ENC_CHECK(grow(e->perm, e->perm_bytes, words * 4 + 16));
    uint32_t* P = (uint32_t*)e->perm;
    if ((G & 31u) == 0) {
        const uint32_t rw = G >> 5;
        if ((rw & 3u) == 0 && (((uintptr_t)lin) & 15u) == 0)
            permute_kernel<4><<<(unsigned)((uint64_t)G * (rw / 4)), 256, 0, st>>>(lin, P, G, order);
        else
            permute_kernel<1><<<(unsigned)((uint64_t)G * rw), 256, 0, st>>>(lin, P, G, order);
    } else {
        permute_gather_kernel<<<grid_for(words), 256, 0, st>>>(lin, P, G, order);
    }
    ENC_HIP(hipGetLastError(), "permute kernel");
    if (format == VOX_ENC_VOX) {
        ENC_CHECK(encode_vox(e, P, G));
        *out_bytes = e->result;
        return VOX_OK;
    }

    const uint64_t nb = (words + kScanTile - 1) / kScanTile;
    ENC_CHECK(grow(e->cnt, e->cnt_bytes, words * 4));
    ENC_CHECK(grow(e->off, e->off_bytes, words * 8));
    ENC_CHECK(grow(e->sums, e->sums_bytes, (nb + 1) * 8));
    uint32_t* cnt = (uint32_t*)e->cnt;
    uint64_t* off = (uint64_t*)e->off;
    uint64_t* sums = (uint64_t*)e->sums;
    const unsigned gb = grid_for(words);

    if (format == VOX_ENC_BINVOX) {
        // tiles of kRleTile words: last change per tile -> exclusive max-scan (one
        // workgroup) -> pairs per tile -> exclusive sum-scan -> pairs emitted
        const uint64_t nt = (words + kRleTile - 1) / kRleTile, nth = nt * 256;
        ENC_CHECK(grow(e->aux, e->aux_bytes, nt * 8 + nth * 20 + 16));
        int64_t* tile_s = (int64_t*)e->aux;                            // last change, then run start, per tile
        int64_t* thread_s = tile_s + nt;                               // run start at each thread's words
        int64_t* thread_first = thread_s + nth;                        // start of each thread's first pair
        uint32_t* thread_off = (uint32_t*)(thread_first + nth);       // pair offset inside the tile
        uint64_t* tile_off = off;                                      // pairs per tile, then their prefix
        rle_tile_last_kernel<<<(unsigned)nt, 256, 0, st>>>(P, words, n, tile_s);
        scan_top_kernel<MaxOp><<<1, 256, 0, st>>>((uint64_t*)tile_s, nt);
        rle_tile_count_kernel<<<(unsigned)nt, 256, 0, st>>>(P, words, n, tile_s, tile_off, thread_s, thread_off,
                                                            thread_first);
        ENC_HIP(hipMemcpyAsync(e->host + 1, tile_off + nt - 1, 8, hipMemcpyDeviceToHost, st), "read-back");
        scan_top_kernel<SumOp><<<1, 256, 0, st>>>(tile_off, nt);
        ENC_HIP(hipMemcpyAsync(e->host, tile_off + nt - 1, 8, hipMemcpyDeviceToHost, st), "read-back");
        // the pairs are emitted into the buffer as it is (a previous call's size, or a quarter
        // pair per word) before their number is known; a larger result is emitted again
        ENC_CHECK(grow(e->out, e->out_bytes, words / 2 + 16));
        uint64_t cap = (e->out_bytes - 16) / 2;
        rle_tile_emit_kernel<<<(unsigned)nt, 256, 0, st>>>(P, words, n, thread_s, thread_off, thread_first, nth,
                                                           tile_off, cap, (uint8_t*)e->out);
        ENC_HIP(hipGetLastError(), "binvox kernels");
        ENC_HIP(hipStreamSynchronize(st), "hipStreamSynchronize");
        const uint64_t np = e->host[0] + e->host[1];
        if (np > cap) {
            ENC_CHECK(grow(e->out, e->out_bytes, np * 2 + 16));
            cap = (e->out_bytes - 16) / 2;
            rle_tile_emit_kernel<<<(unsigned)nt, 256, 0, st>>>(P, words, n, thread_s, thread_off, thread_first, nth,
                                                               tile_off, cap, (uint8_t*)e->out);
            ENC_HIP(hipGetLastError(), "binvox emit kernel");
        }
        e->result = np * 2;
    } else {
        obj_kernel<false><<<gb, 256, 0, st>>>(P, words, G, format, cnt, nullptr, nullptr);
        ENC_HIP((device_scan<SumOp, uint32_t>(cnt, words, sums, off, st)), "sum-scan");
        e->host[1] = 0;
        ENC_HIP(hipMemcpyAsync(e->host, off + words - 1, 8, hipMemcpyDeviceToHost, st), "read-back");
        ENC_HIP(hipMemcpyAsync(e->host + 1, cnt + words - 1, 4, hipMemcpyDeviceToHost, st), "read-back");
        ENC_HIP(hipStreamSynchronize(st), "hipStreamSynchronize");
        const uint64_t total = e->host[0] + (e->host[1] & 0xFFFFFFFFull);
        ENC_CHECK(grow(e->out, e->out_bytes, total + 16));
        obj_kernel<true><<<gb, 256, 0, st>>>(P, words, G, format, nullptr, off, (uint8_t*)e->out);
        ENC_HIP(hipGetLastError(), "obj kernels");
        e->result = total;
    }
    *out_bytes = e->result;
    return VOX_OK;
}

vox_status vox_encoder_data(vox_encoder* e, const uint8_t** d_bytes) {
    if (!e || !d_bytes) return fail(VOX_ERR_INVALID, "NULL argument");
    *d_bytes = (const uint8_t*)e->out;
    return VOX_OK;
}

vox_status vox_encoder_copy(vox_encoder* e, void* host, uint64_t bytes) {
    if (!e || (bytes && !host)) return fail(VOX_ERR_INVALID, "NULL argument");
    if (bytes > e->result) return fail(VOX_ERR_INVALID, "more bytes requested than the last encoding produced");
    if (bytes == 0) return VOX_OK;
    ENC_HIP(hipMemcpyAsync(host, e->out, bytes, hipMemcpyDeviceToHost, e->stream), "hipMemcpyAsync");
    ENC_HIP(hipStreamSynchronize(e->stream), "hipStreamSynchronize");
    return VOX_OK;
}

int vox_binvox_header(const vox_info* info, char* buf, size_t cap) {
    if (!info) return -1;
    const float sx = info->bbox_max[0] - info->bbox_min[0], sy = info->bbox_max[1] - info->bbox_min[1],
                sz = info->bbox_max[2] - info->bbox_min[2];
    // std::max(std::max(a, b), c) returns its first argument unless the second is larger
    float s = sx < sy ? sy : sx;
    s = s < sz ? sz : s;
    const int len = snprintf(buf, cap, "#binvox 1\ndim %u %u %u\ntranslate %g %g %g\nscale %g\ndata\n",
                             info->gridsize[0], info->gridsize[1], info->gridsize[2], (double)info->bbox_min[0],
                             (double)info->bbox_min[1], (double)info->bbox_min[2], (double)s);
    return len;
}

}  // extern "C"
