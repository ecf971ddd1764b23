// raster.h -- wave-balanced surface rasterization of 64 triangles at a time.
//
// A lane that owns one triangle would walk its own z/y/x voxel box, and a
// wave would run as long as its largest box.  Instead every lane publishes
// its triangle's test coefficients to LDS, the wave takes an exclusive scan
// of the per-triangle work items -- (y, z) rows split into 32-voxel x
// chunks -- and deals the items round-robin over its 64 lanes.  Each item
// runs the YZ test once, then the plane, XY and ZX tests for each voxel of
// its x chunk (vox_device.h arithmetic), and reports the passing voxels as a
// bit mask to the caller's emitter.
#pragma once

#include "vox_device.h"

#pragma clang fp contract(off)

namespace voxamd {

// Low 10 bits of v spread to every third bit (bits 0, 3, ..., 27).
__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// Coefficients a row needs, as published to LDS (38 words).
struct SurfRec {
    float n[3], d1, d2;
    float xy[9], zx[9], yz[9];  // a[3], b[3], d[3] per projection
    int xa, xb, ya, yb, za, zb;  // voxel box to cover (inclusive, possibly clipped)
};

struct WaveRaster {
    SurfRec rec[64];
    unsigned incl[64];  // inclusive prefix of item counts
};

__device__ __forceinline__ void publish(SurfRec& R, const SurfTri& T, int xa, int xb, int ya, int yb, int za,
                                        int zb) {
#pragma unroll
    for (int k = 0; k < 3; k++) R.n[k] = T.n[k];
    R.d1 = T.d1;
    R.d2 = T.d2;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        R.xy[i] = T.xy.a[i]; R.xy[3 + i] = T.xy.b[i]; R.xy[6 + i] = T.xy.d[i];
        R.zx[i] = T.zx.a[i]; R.zx[3 + i] = T.zx.b[i]; R.zx[6 + i] = T.zx.d[i];
        R.yz[i] = T.yz.a[i]; R.yz[3 + i] = T.yz.b[i]; R.yz[6 + i] = T.yz.d[i];
    }
    R.xa = xa; R.xb = xb; R.ya = ya; R.yb = yb; R.za = za; R.zb = zb;
}

__device__ __forceinline__ bool pass3(const float* E, float pu, float pv) {
    bool ok = true;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        float t = E[i] * pu;
        t = t + E[3 + i] * pv;
        t = t + E[6 + i];
        ok = ok && !(t < 0.0f);
    }
    return ok;
}

// Items of a box: rows x 32-voxel chunks aligned to multiples of 32 in x.
__device__ __forceinline__ unsigned box_items(int xa, int xb, int ya, int yb, int za, int zb) {
    if (xa > xb || ya > yb || za > zb) return 0u;
    const unsigned nch = (unsigned)((xb >> 5) - (xa >> 5) + 1);
    return (unsigned)(yb - ya + 1) * (unsigned)(zb - za + 1) * nch;
}

// Rasterize the wave's triangles.  `have` says whether this lane brings a
// triangle; (xa..zb) is its (clipped) voxel box.  emit(y, z, cx, mask) gets
// bit k of mask set for each passing voxel cx + k (k < 32, cx a multiple of
// 32).  W is this wave's LDS area.  Every wave of the workgroup must call it
// the same number of times (it synchronises the workgroup).
template <class Emit>
__device__ __forceinline__ void wave_raster(WaveRaster& W, const vox_info& info, bool have, const SurfTri& T,
                                            int xa, int xb, int ya, int yb, int za, int zb, Emit&& emit) {
    const int lane = __lane_id();
    unsigned items = have ? box_items(xa, xb, ya, yb, za, zb) : 0u;
    if (items) publish(W.rec[lane], T, xa, xb, ya, yb, za, zb);
    // inclusive scan of item counts over the wave
    unsigned incl = items;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned o = __shfl_up(incl, d);
        if (lane >= d) incl += o;
    }
    W.incl[lane] = incl;
    const unsigned total = __shfl(incl, 63);
    __syncthreads();
    const float ux = info.unit[0], uy = info.unit[1], uz = info.unit[2];
    for (unsigned j = lane; j < total; j += 64) {
        // owner: first lane whose inclusive prefix exceeds j
        int lo = 0, hi = 63;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (W.incl[mid] > j) hi = mid; else lo = mid + 1;
        }
        const SurfRec& R = W.rec[lo];
        const unsigned q = j - (lo ? W.incl[lo - 1] : 0u);
        const int c0 = R.xa >> 5;
        const unsigned nch = (unsigned)((R.xb >> 5) - c0 + 1);
        const unsigned ny = (unsigned)(R.yb - R.ya + 1);
        const unsigned row = q / nch, ch = q - row * nch;
        const int y = R.ya + (int)(row % ny), z = R.za + (int)(row / ny);
        const float py = (float)y * uy, pz = (float)z * uz;
        if (!pass3(R.yz, py, pz)) continue;
        const int cx = (c0 + (int)ch) << 5;
        const int x0 = cx > R.xa ? cx : R.xa, x1 = (cx + 31) < R.xb ? (cx + 31) : R.xb;
        // row-invariant parts of the plane, XY and ZX tests
        const float nypy = R.n[1] * py, nzpz = R.n[2] * pz;
        float bxy[3], azx[3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            bxy[i] = R.xy[3 + i] * py;
            azx[i] = R.zx[i] * pz;
        }
        uint32_t mask = 0;
        for (int x = x0; x <= x1; x++) {
            const float px = (float)x * ux;
            float t = R.n[0] * px;
            t = t + nypy;
            t = t + nzpz;
            const float r = (t + R.d1) * (t + R.d2);
            bool ok = !(r > 0.0f);
#pragma unroll
            for (int i = 0; i < 3; i++) {
                float a = R.xy[i] * px;
                a = a + bxy[i];
                a = a + R.xy[6 + i];
                float b = azx[i];
                b = b + R.zx[3 + i] * px;
                b = b + R.zx[6 + i];
                ok = ok && !(a < 0.0f) && !(b < 0.0f);
            }
            if (ok) mask |= 1u << (x - cx);
        }
        if (mask) emit(y, z, cx, mask);
    }
    __syncthreads();
}

// Split a 32-voxel x-chunk mask (bit k = voxel cx + k, cx a multiple of 32) of
// row (y, z) into Morton table words: x bits 0 and 1 land in code bits 0 and 3,
// so every group of 4 consecutive x shares one word.  fn(word, bits).
template <class Fn>
__device__ __forceinline__ void morton_chunk_words(uint64_t base_code, uint32_t yz_low, uint32_t mask, Fn&& fn) {
    // yz_low: code bits 1, 2, 4 contributed by y0, z0, y1 (a value in 0..31)
    while (mask) {
        const int g = (__ffs(mask) - 1) >> 2;  // group of 4 x values
        const uint32_t nib = (mask >> (4 * g)) & 0xFu;
        mask &= ~(0xFu << (4 * g));
        const uint64_t code = base_code | (uint64_t)spread10((uint32_t)(4 * g));
        uint32_t bits = 0;
        const uint32_t j0 = yz_low;  // j for x offset 0 in the group
        if (nib & 1u) bits |= 1u << (31u - j0);
        if (nib & 2u) bits |= 1u << (31u - (j0 | 1u));
        if (nib & 4u) bits |= 1u << (31u - (j0 | 8u));
        if (nib & 8u) bits |= 1u << (31u - (j0 | 9u));
        fn(code >> 5, bits);
    }
}

}  // namespace voxamd
