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

// Coefficients a row needs, as published to LDS (42 words).
struct SurfRec {
    float n[3], d1, d2;
    float xy[9], zx[9], yz[9];  // a[3], b[3], d[3] per projection
    int xa, xb, ya, yb, za, zb;  // voxel box to cover (inclusive, possibly clipped)
    int nk;                      // x chunks enumerated per row (see plane_band)
    int band;                    // 1: prune x with the plane band
    float sx;                    // |n.x| * (xb + 1) * unit.x, the x part of the magnitude scale
    int pad;
};

// Conservative x range of a row (y, z) outside of which the plane test
// (n.p + d1) * (n.p + d2) > 0 certainly rejects.  In exact arithmetic the test
// passes iff T = n.x*x*ux + C lies in [-max(d1,d2), -min(d1,d2)] (C = n.y*py
// + n.z*pz; the sign of a float sum is the sign of the exact sum, and a
// product can only underflow to zero within ~1e-41 of the band edge).  The
// float evaluation of T is within a few ulp of the magnitudes involved; the
// margin allows 64 ulp plus two voxels on each side.  Every voxel inside the
// range is still decided by the exact float tests.
__device__ __forceinline__ void plane_band(const SurfRec& R, float ux, float py, float pz, int& lo, int& hi) {
    const double s = (double)R.n[0] * (double)ux;
    const double C = (double)R.n[1] * (double)py + (double)R.n[2] * (double)pz;
    const double dmax = R.d1 > R.d2 ? (double)R.d1 : (double)R.d2;
    const double dmin = R.d1 > R.d2 ? (double)R.d2 : (double)R.d1;
    const double S = (double)R.sx + fabs((double)R.n[1] * (double)py) + fabs((double)R.n[2] * (double)pz) +
                     fabs(dmax) + fabs(dmin);
    const double M = 64.0 * 5.9604644775390625e-08 * S + 1e-30;
    double a = (-dmax - M - C) / s, b = (-dmin + M - C) / s;
    if (a > b) {
        const double t = a;
        a = b;
        b = t;
    }
    lo = R.xa;
    hi = R.xb;
    if (!(a <= b)) return;  // not a number: no pruning
    // every compare below happens in double before any conversion to int
    if (a > (double)R.xb) { lo = R.xb + 1; return; }  // band entirely right of the box
    if (b < (double)R.xa) { hi = R.xa - 1; return; }  // entirely left
    if (a > (double)R.xa) lo = (int)floor(a) - 2;
    if (b < (double)R.xb) hi = (int)ceil(b) + 2;
    if (lo < R.xa) lo = R.xa;
    if (hi > R.xb) hi = R.xb;
}

struct WaveRaster {
    SurfRec rec[64];
    unsigned incl[64];  // inclusive prefix of item counts
};

__device__ __forceinline__ void publish(SurfRec& R, const SurfTri& T, int xa, int xb, int ya, int yb, int za,
                                        int zb, int nk, int band, float sx) {
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
    R.nk = nk;
    R.band = band;
    R.sx = sx;
}

// Chunks per row to enumerate: all x chunks of the box, or -- when the plane
// band is narrow -- only as many as the widest band of any row can touch.
__device__ __forceinline__ void band_chunks(const SurfTri& T, const vox_info& info, int xa, int xb, int yb, int zb,
                                            int& nk, int& band, float& sx) {
    const int nch = (xb >> 5) - (xa >> 5) + 1;
    nk = nch;
    band = 0;
    sx = 0.0f;
    const float ax = fabsf(T.n[0]);
    if (!(ax > 0.0f) || !isfinite(T.n[1]) || !isfinite(T.n[2]) || !isfinite(T.d1) || !isfinite(T.d2)) return;
    sx = ax * (float)(xb + 1) * info.unit[0] * 1.0001f;
    const double s = (double)ax * (double)info.unit[0];
    const double Smax = (double)sx + fabs((double)T.n[1]) * (yb + 1.0) * info.unit[1] +
                        fabs((double)T.n[2]) * (zb + 1.0) * info.unit[2] + fabs((double)T.d1) + fabs((double)T.d2);
    const double M = 64.0 * 5.9604644775390625e-08 * Smax + 1e-30;
    const double width = (fabs((double)T.d1 - (double)T.d2) + 2.0 * M) / s + 6.0;  // voxels, with the +-2 slack
    if (!(width < 1e9)) return;
    const int k = (int)ceil(width / 32.0) + 1;
    if (k < nch) {
        nk = k;
        band = 1;
    } else {
        band = 1;  // still prune x inside the enumerated chunks
    }
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

// Items of a box: rows x (enumerated) 32-voxel chunks.
__device__ __forceinline__ unsigned box_items(int xa, int xb, int ya, int yb, int za, int zb, int nk) {
    if (xa > xb || ya > yb || za > zb) return 0u;
    return (unsigned)(yb - ya + 1) * (unsigned)(zb - za + 1) * (unsigned)nk;
}

// Mask of the passing voxels x0..x1 (inside the 32-voxel chunk at cx) of row
// (y, z) = (py, pz) for coefficients (n, d1, d2, xy[9], zx[9]); the per-voxel
// arithmetic of vox_device.h's plane_pass / edges_pass with the row-invariant
// products hoisted (each product and sum is the same binary32 operation).
__device__ __forceinline__ uint32_t row_mask(const float* n, float d1, float d2, const float* xy, const float* zx,
                                             float ux, float py, float pz, int x0, int x1, int cx) {
    const float nypy = n[1] * py, nzpz = n[2] * pz;
    float bxy[3], azx[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        bxy[i] = xy[3 + i] * py;
        azx[i] = zx[i] * pz;
    }
    uint32_t mask = 0;
    for (int x = x0; x <= x1; x++) {
        const float px = (float)x * ux;
        float t = n[0] * px;
        t = t + nypy;
        t = t + nzpz;
        const float r = (t + d1) * (t + d2);
        bool ok = !(r > 0.0f);
#pragma unroll
        for (int i = 0; i < 3; i++) {
            float a = xy[i] * px;
            a = a + bxy[i];
            a = a + xy[6 + i];
            float b = azx[i];
            b = b + zx[3 + i] * px;
            b = b + zx[6 + i];
            ok = ok && !(a < 0.0f) && !(b < 0.0f);
        }
        if (ok) mask |= 1u << (x - cx);
    }
    return mask;
}

// Boxes this small are rasterized by their own lane (no LDS round trip, no
// work-item search); larger ones are shared over the wave.
constexpr int kLaneX = 32;      // x extent (at most two 32-voxel chunks)
constexpr int kLaneRows = 16;   // (y, z) rows
constexpr int kLaneVox = 64;    // voxels

// Wave-level barrier: the wave's LDS writes are visible to all its lanes.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Rasterize the wave's triangles.  `have` says whether this lane brings a
// triangle; (xa..zb) is its (clipped) voxel box.  emit(y, z, cx, mask) gets
// bit k of mask set for each passing voxel cx + k (k < 32, cx a multiple of
// 32).  W is this wave's LDS area (synchronised at wave level only).
template <class Emit>
__device__ __forceinline__ void wave_raster(WaveRaster& W, const vox_info& info, bool have, const SurfTri& T,
                                            int xa, int xb, int ya, int yb, int za, int zb, Emit&& emit) {
    const int lane = __lane_id();
    const float ux = info.unit[0], uy = info.unit[1], uz = info.unit[2];
    const bool empty = !have || xa > xb || ya > yb || za > zb;
    const int nx = xb - xa + 1, nrows = (yb - ya + 1) * (zb - za + 1);
    const bool small = !empty && nx <= kLaneX && nrows <= kLaneRows && nx * nrows <= kLaneVox;
    if (small) {
        const float xy[9] = {T.xy.a[0], T.xy.a[1], T.xy.a[2], T.xy.b[0], T.xy.b[1], T.xy.b[2],
                             T.xy.d[0], T.xy.d[1], T.xy.d[2]};
        const float zx[9] = {T.zx.a[0], T.zx.a[1], T.zx.a[2], T.zx.b[0], T.zx.b[1], T.zx.b[2],
                             T.zx.d[0], T.zx.d[1], T.zx.d[2]};
        const int c0 = xa & ~31, c1 = xb & ~31;
        for (int z = za; z <= zb; z++) {
            const float pz = (float)z * uz;
            for (int y = ya; y <= yb; y++) {
                const float py = (float)y * uy;
                if (!edges_pass(T.yz, py, pz)) continue;
                const uint32_t m0 = row_mask(T.n, T.d1, T.d2, xy, zx, ux, py, pz, xa, c1 > c0 ? c0 + 31 : xb, c0);
                if (m0) emit(y, z, c0, m0);
                if (c1 > c0) {
                    const uint32_t m1 = row_mask(T.n, T.d1, T.d2, xy, zx, ux, py, pz, c1, xb, c1);
                    if (m1) emit(y, z, c1, m1);
                }
            }
        }
    }
    const bool big = !empty && !small;
    if (!__ballot(big)) return;  // uniform over the wave
    int nk = 1, band = 0;
    float sx = 0.0f;
    if (big) band_chunks(T, info, xa, xb, yb, zb, nk, band, sx);
    unsigned items = big ? box_items(xa, xb, ya, yb, za, zb, nk) : 0u;
    if (items) publish(W.rec[lane], T, xa, xb, ya, yb, za, zb, nk, band, sx);
    // inclusive scan of item counts over the wave
    unsigned incl = items;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned o = __shfl_up(incl, d);
        if (lane >= d) incl += o;
    }
    W.incl[lane] = incl;
    const unsigned total = __shfl(incl, 63);
    wave_sync();
    for (unsigned j = lane; j < total; j += 64) {
        // owner: first lane whose inclusive prefix exceeds j
        int lo = 0, hi = 63;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (W.incl[mid] > j) hi = mid; else lo = mid + 1;
        }
        const SurfRec& R = W.rec[lo];
        const unsigned q = j - (lo ? W.incl[lo - 1] : 0u);
        const unsigned nkr = (unsigned)R.nk;
        const unsigned ny = (unsigned)(R.yb - R.ya + 1);
        const unsigned row = q / nkr, k = q - row * nkr;
        const int y = R.ya + (int)(row % ny), z = R.za + (int)(row / ny);
        const float py = (float)y * uy, pz = (float)z * uz;
        int bx0 = R.xa, bx1 = R.xb;
        if (R.band) {
            plane_band(R, ux, py, pz, bx0, bx1);
            if (bx0 > bx1) continue;
        }
        const int cx = ((bx0 >> 5) + (int)k) << 5;
        if (cx > bx1) continue;
        if (!pass3(R.yz, py, pz)) continue;
        const int x0 = cx > bx0 ? cx : bx0, x1 = (cx + 31) < bx1 ? (cx + 31) : bx1;
        const uint32_t mask = row_mask(R.n, R.d1, R.d2, R.xy, R.zx, ux, py, pz, x0, x1, cx);
        if (mask) emit(y, z, cx, mask);
    }
    wave_sync();  // W is reused by the wave's next call
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
        // spread10(4 g) for g < 8: x bits 2, 3, 4 land in code bits 6, 9, 12
        const uint64_t code = base_code | (uint64_t)((((uint32_t)g & 1u) << 6) | (((uint32_t)g & 2u) << 8) |
                                                     (((uint32_t)g & 4u) << 10));
        // x offsets 0..3 of the group are code bits 0 and 3 (j0 | 0, 1, 8, 9; yz_low
        // never has bits 0 or 3), i.e. word bits 31, 30, 23, 22 shifted down by j0
        const uint32_t bits = (((nib & 1u) << 31) | ((nib & 2u) << 29) | ((nib & 4u) << 21) | ((nib & 8u) << 19)) >>
                              yz_low;
        fn(code >> 5, bits);
    }
}

}  // namespace voxamd
