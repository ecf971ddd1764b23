// bricks.h -- LDS row blocks of a solid table: the suffix XOR along x and the
// write-out of 32 y x rz z rows as Morton bricks.
//
// A block holds rows (y0 + y, z0 + z) for y < 32, z < rz, all x: row r = z * 32
// + y at tile[r * ts], rw = G / 32 words each (ts = rw + 1: the pad word
// spreads a column of rows over the LDS banks).  A row of run-end markers
// becomes the solid row by its suffix XOR along x (bit x = parity of the
// markers at x' >= x; larger x = lower bit of a word, later words).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vox_device.h"

namespace voxamd {

// In-word suffix along x: bit b (x = 31 - b) becomes the XOR of bits 0..b.
__device__ __forceinline__ uint32_t prefix_xor_up(uint32_t w) {
    w ^= w << 1;
    w ^= w << 2;
    w ^= w << 4;
    w ^= w << 8;
    w ^= w << 16;
    return w;
}

// Suffix XOR of rows 0..nrows-1 of tile (stride ts, rw words per row) by a
// workgroup of four waves; the final word j of row r goes to store(r, j, w).
// rw <= 64: a wave holds 64 / rw whole rows (lane = word) and reads four
// row groups before storing any; rw > 64: a wave walks its row from the last
// 64-word segment down, carrying the parity.  Rows are disjoint between
// waves, so store may write the result back into the tile.
template <class Store>
__device__ __forceinline__ void tile_suffix(const uint32_t* tile, uint32_t ts, uint32_t rw, uint32_t nrows,
                                            Store&& store) {
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
    if (rw <= 64) {
        const uint32_t per = 64 / rw;
        const uint32_t r_in = (uint32_t)lane / rw, j = (uint32_t)lane % rw;
        const bool active_lane = r_in < per;
        for (uint32_t r00 = (uint32_t)wave * per; r00 < nrows; r00 += 16 * per) {
            uint32_t w[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const uint32_t r = r00 + 4 * per * (uint32_t)u + r_in;
                w[u] = active_lane && r < nrows ? tile[r * ts + j] : 0u;
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const uint32_t r = r00 + 4 * per * (uint32_t)u + r_in;
                const bool act = active_lane && r < nrows;
                const unsigned long long par = __ballot(act && (__popc(w[u]) & 1));
                const uint32_t seg0 = r_in * rw;
                const unsigned long long above =
                    (active_lane && j + 1 < rw) ? ((par >> (seg0 + j + 1)) & ((1ull << (rw - 1 - j)) - 1ull))
                                                : 0ull;
                const uint32_t fin = prefix_xor_up(w[u]) ^ ((__popcll(above) & 1) ? 0xFFFFFFFFu : 0u);
                if (act) store(r, j, fin);
            }
        }
    } else {
        for (uint32_t r = (uint32_t)wave; r < nrows; r += 4) {
            uint32_t carry = 0;
            for (int64_t hi = (int64_t)rw; hi > 0; hi -= 64) {
                const int64_t lo = hi > 64 ? hi - 64 : 0;
                const bool act = (int64_t)lane < hi - lo;
                const uint32_t j = (uint32_t)lo + (uint32_t)lane;
                const uint32_t w = act ? tile[r * ts + j] : 0u;
                const unsigned long long par = __ballot(act && (__popc(w) & 1));
                const unsigned long long above = lane < 63 ? (par >> (lane + 1)) : 0ull;
                const uint32_t fin = prefix_xor_up(w) ^ (((__popcll(above) + carry) & 1) ? 0xFFFFFFFFu : 0u);
                carry = (uint32_t)((__popcll(par) + carry) & 1);
                if (act) store(r, j, fin);
            }
        }
    }
}

// The block's rows (a power-of-two G) -> Morton table words: a word is a
// 4x4x2 brick (code bits 0..4 = x0 y0 z0 x1 y1).  Stored, or XORed into the
// table when accumulating.  Every thread of the workgroup (256) takes part.
__device__ __forceinline__ void tile_to_bricks(const uint32_t* tile, uint32_t ts, uint32_t* __restrict__ table,
                                               uint32_t G, uint32_t rw, uint32_t rz, uint32_t y0, uint32_t z0,
                                               int accumulate) {
    if (rz == 8u && (((uintptr_t)table) & 15u) == 0) {
        // 32 y x 8 z rows: the block's Morton words, with the code bits above
        // bit 5 fixed by (y0, z0), vary in code bits 5..10 (z1 x2 y2 z2 x3 y3),
        // 12, 13 (x4 y4) and 3k for x_k, k >= 5.  A thread takes 8 words that
        // are consecutive in memory (code bits 5..7 = z1 x2 y2) and stores them
        // as two 16-byte writes; its index supplies z2 x3 y3 x4 y4 x5.. , so
        // eight neighbouring lanes write 256 contiguous bytes.  Its 8 words
        // draw on 32 rows (y0 y1 y2 z0 z1 free) and one byte (8 x) of each.
        const uint64_t base_w = morton3(0u, y0, z0) >> 5;
        for (uint32_t t = threadIdx.x; t < G; t += 256) {
            const uint32_t z2 = t & 1u, x3 = (t >> 1) & 1u, y3 = (t >> 2) & 1u, x4 = (t >> 3) & 1u,
                           y4 = (t >> 4) & 1u, xh = t >> 5;
            const uint32_t xbyte = x3 | (x4 << 1) | (xh << 2);  // x >> 3
            const uint32_t j = xbyte >> 2, bsh = 24u - 8u * (xbyte & 3u);
            uint32_t bytes[32];  // row (y0 y1 y2 z0 z1) -> its byte of x
#pragma unroll
            for (int q = 0; q < 32; q++) {
                const uint32_t y_l = ((uint32_t)q & 7u) | (y3 << 3) | (y4 << 4);
                const uint32_t z_l = (((uint32_t)q >> 3) & 3u) | (z2 << 2);
                bytes[q] = (tile[(z_l * 32u + y_l) * ts + j] >> bsh) & 0xFFu;
            }
            // The 256 bits (row q = y0 y1 y2 z0 z1, x offset xo = x0 x1 x2 within the byte, MSB
            // first) become the 8 words k = z1 | x2 << 1 | y2 << 2, bit 31 - c with
            // c = x0 | y0 << 1 | z0 << 2 | x1 << 3 | y1 << 4.  Word w = y2 | z0 << 1 | z1 << 2 packs
            // the bytes of rows (y0, y1) = (0,0) (1,0) (0,1) (1,1) from the top, so bit p of it
            // has p0 = !x0, p1 = !x1, p2 = !x2, p3 = !y0, p4 = !y1.  Exchanging the low nibbles
            // of word pair (z0 = 0, 1) with the high ones puts x2 in the word index (z0 = 0
            // keeps x2 = 0) and !z0 at p2; swapping index bits p1 and p3 (a delta swap) leaves
            // p = 31 - c.
            uint32_t W[8];
#pragma unroll
            for (int w = 0; w < 8; w++)
                W[w] = (bytes[4 * w] << 24) | (bytes[4 * w + 1] << 16) | (bytes[4 * w + 2] << 8) | bytes[4 * w + 3];
#pragma unroll
            for (int w = 0; w < 8; w++) {
                if (w & 2) continue;
                const uint32_t t = (W[w] ^ (W[w | 2] >> 4)) & 0x0F0F0F0Fu;
                W[w] ^= t;
                W[w | 2] ^= t << 4;
            }
            uint32_t out[8];
#pragma unroll
            for (int w = 0; w < 8; w++) {
                const uint32_t t = ((W[w] >> 6) ^ W[w]) & 0x00CC00CCu;
                out[((w >> 2) & 1) | (((w >> 1) & 1) << 1) | ((w & 1) << 2)] = W[w] ^ t ^ (t << 6);
            }
            // word offset inside the block's code range: bits 3 4 5 7 8 from t, then x5.. at 10, 13, ...
            uint64_t wl = ((uint64_t)z2 << 3) | ((uint64_t)x3 << 4) | ((uint64_t)y3 << 5) | ((uint64_t)x4 << 7) |
                          ((uint64_t)y4 << 8);
            for (uint32_t b = 0; (xh >> b) != 0u; b++) wl |= (uint64_t)((xh >> b) & 1u) << (10u + 3u * b);
            uint4* dst = (uint4*)(table + base_w + wl);
            uint4 lo = make_uint4(out[0], out[1], out[2], out[3]), hi = make_uint4(out[4], out[5], out[6], out[7]);
            if (accumulate) {
                const uint4 a = dst[0], b = dst[1];
                lo.x ^= a.x; lo.y ^= a.y; lo.z ^= a.z; lo.w ^= a.w;
                hi.x ^= b.x; hi.y ^= b.y; hi.z ^= b.z; hi.w ^= b.w;
            }
            dst[0] = lo;
            dst[1] = hi;
        }
        return;
    }
    const uint32_t groups = 8u * (rz / 2u);  // brick rows (4 y x 2 z) in the block
    for (uint32_t t = threadIdx.x; t < groups * rw; t += 256) {
        const uint32_t j = t % rw, br = t / rw;
        const uint32_t yb = br % 8u, zb = br / 8u;
        uint32_t row_w[8];  // row q = (y1 << 2 | z0 << 1 | y0) -> its linear word j
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const uint32_t y0b = (uint32_t)q & 1u, z0b = ((uint32_t)q >> 1) & 1u, y1b = ((uint32_t)q >> 2) & 1u;
            row_w[q] = tile[((zb * 2u + z0b) * 32u + yb * 4u + y0b + 2u * y1b) * ts + j];
        }
#pragma unroll
        for (int xg = 0; xg < 8; xg++) {
            uint32_t word = 0;
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const uint32_t y0b = (uint32_t)q & 1u, z0b = ((uint32_t)q >> 1) & 1u, y1b = ((uint32_t)q >> 2) & 1u;
                // nibble bits 3..0 = x offsets 0..3; offset o = x0 + 2 x1 goes to code bit c0 | x0 | x1 << 3,
                // i.e. word bit 31 - c0 - x0 - 8 x1: offsets 0, 1 are the adjacent word bits 31-c0, 30-c0
                // and offsets 2, 3 the adjacent bits 23-c0, 22-c0 (c0 is a compile-time constant here)
                const uint32_t nib = (row_w[q] >> (28 - 4 * xg)) & 0xFu;
                const uint32_t c0 = (y0b << 1) | (z0b << 2) | (y1b << 4);
                word |= ((nib >> 2) << (30u - c0)) | ((nib & 3u) << (22u - c0));
            }
            const uint64_t code = morton3(j * 32u + (uint32_t)xg * 4u, y0 + yb * 4u, z0 + zb * 2u);
            uint32_t* dst = table + (code >> 5);
            *dst = accumulate ? (*dst ^ word) : word;
        }
    }
}

}  // namespace voxamd
