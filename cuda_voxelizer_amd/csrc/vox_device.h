// vox_device.h -- device-side geometry shared by every voxelization kernel.
//
// Bit parity contract: every binary32 operation below is the one the
// reference CPU voxelizer evaluates (cpu_voxelizer.cpp:56-171 surface,
// :255-310 solid, with the host semantics of helper_math.h:51-81), in the same
// association order, with FP contraction off (the pragma below, and
// -ffp-contract=off on the command line), IEEE f32 denormals and correctly
// rounded division and sqrt (hipcc defaults).  The kernels are free to
// reorganise *which* voxels they evaluate and *when*; the value of each
// individual test is fixed by this file.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cuda_voxelizer_amd/vox_capi.h"

#pragma clang fp contract(off)

namespace voxamd {

// reference float_error (cpu_voxelizer.cpp:2) is the double literal 0.000001.
constexpr double kFloatError = 0.000001;

__device__ __forceinline__ float hmin(float a, float b) { return a < b ? a : b; }  // host fminf
__device__ __forceinline__ float hmax(float a, float b) { return a > b ? a : b; }  // host fmaxf
// std::max<float>(0.0f, a): (0.0f < a) ? a : 0.0f  (timer.h:13 brings std:: into scope)
__device__ __forceinline__ float max0(float a) { return (0.0f < a) ? a : 0.0f; }
__device__ __forceinline__ int iclamp(int f, int lo, int hi) {
    int m = f < hi ? f : hi;
    return lo > m ? lo : m;
}

// One projection (XY, YZ or ZX) of the Schwarz-Seidel overlap test: three
// edge half-planes a*u + b*v + d >= 0 in that projection's (u, v) coords.
struct EdgeTest {
    float a[3], b[3], d[3];
};

// Per-triangle constants of the surface test.
struct SurfTri {
    int lo[3], hi[3];  // clamped voxel bbox (inclusive)
    float n[3];        // unit normal (NaN for zero-area triangles)
    float d1, d2;      // plane-test offsets
    EdgeTest xy, yz, zx;
};

// Copies n <= 9 * NT floats (the vertices of up to NT triangles) from memory to
// LDS with NT threads: each thread issues all its loads before its first LDS
// store (a plain strided loop waits for every load in turn).
template <int NT>
__device__ __forceinline__ void stage_floats(float* dst, const float* __restrict__ src, int n) {
    float v[9];
#pragma unroll
    for (int u = 0; u < 9; u++) {
        const int i = (int)threadIdx.x + u * NT;
        v[u] = i < n ? src[i] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < 9; u++) {
        const int i = (int)threadIdx.x + u * NT;
        if (i < n) dst[i] = v[u];
    }
}

__device__ __forceinline__ void load_tri(const float* __restrict__ tris, uint64_t t, const vox_info& info,
                                         float v[3][3]) {
    const float* p = tris + t * 9;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        v[k][0] = p[3 * k + 0] - info.bbox_min[0];
        v[k][1] = p[3 * k + 1] - info.bbox_min[1];
        v[k][2] = p[3 * k + 2] - info.bbox_min[2];
    }
}

// unit normal = cross(e0, e1) * (1.0f / sqrtf(dot)) ; helper_math.h:1325-1329 host form.
__device__ __forceinline__ void unit_normal(const float e0[3], const float e1[3], float n[3]) {
    float cx = e0[1] * e1[2] - e0[2] * e1[1];
    float cy = e0[2] * e1[0] - e0[0] * e1[2];
    float cz = e0[0] * e1[1] - e0[1] * e1[0];
    float s = cx * cx;
    s = s + cy * cy;
    s = s + cz * cz;
    float inv = 1.0f / sqrtf(s);
    n[0] = cx * inv;
    n[1] = cy * inv;
    n[2] = cz * inv;
}

// Edge half-planes of one projection.  u/w select the projected coordinates
// (normal of edge e is (-e[w], e[u]), cpu_voxelizer.cpp:91-125), `flip`
// negates all three, ua/ub are the unit lengths multiplying a and b in the
// critical-point offset (for ZX the reference uses unit.x with a and unit.z
// with b, :123-125, which is kept as written).
__device__ __forceinline__ void edge_test(EdgeTest& E, const float v[3][3], const float e[3][3], int u, int w,
                                          bool flip, float ua, float ub) {
#pragma unroll
    for (int i = 0; i < 3; i++) {
        float a = -1.0f * e[i][w];
        float b = e[i][u];
        if (flip) {
            a = -a;
            b = -b;
        }
        float dp = a * v[i][u];
        dp = dp + b * v[i][w];
        float d = -1.0f * dp;
        d = d + max0(ua * a);
        d = d + max0(ub * b);
        E.a[i] = a;
        E.b[i] = b;
        E.d[i] = d;
    }
}

// Clamped voxel box of a translated triangle (cpu_voxelizer.cpp:74-80).
__device__ __forceinline__ void surface_box(const float v[3][3], const vox_info& info, int lo[3], int hi[3]) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int gmax = (int)info.gridsize[k] - 1;
        float mn = hmin(v[0][k], hmin(v[1][k], v[2][k]));
        float mx = hmax(v[0][k], hmax(v[1][k], v[2][k]));
        lo[k] = iclamp((int)(mn / info.unit[k]), 0, gmax);
        hi[k] = iclamp((int)(mx / info.unit[k]), 0, gmax);
    }
}

// Everything of surface_setup except the box (T.lo / T.hi are left alone).
__device__ __forceinline__ void surface_setup_tests(const float v[3][3], const vox_info& info, SurfTri& T) {
    float e[3][3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        e[0][k] = v[1][k] - v[0][k];
        e[1][k] = v[2][k] - v[1][k];
        e[2][k] = v[0][k] - v[2][k];
    }
    unit_normal(e[0], e[1], T.n);
    // critical point c and the two plane offsets (cpu_voxelizer.cpp:83-87)
    float c[3];
#pragma unroll
    for (int k = 0; k < 3; k++) c[k] = (T.n[k] > 0.0f) ? info.unit[k] : 0.0f;
    float d1 = T.n[0] * (c[0] - v[0][0]);
    d1 = d1 + T.n[1] * (c[1] - v[0][1]);
    d1 = d1 + T.n[2] * (c[2] - v[0][2]);
    float d2 = T.n[0] * ((info.unit[0] - c[0]) - v[0][0]);
    d2 = d2 + T.n[1] * ((info.unit[1] - c[1]) - v[0][1]);
    d2 = d2 + T.n[2] * ((info.unit[2] - c[2]) - v[0][2]);
    T.d1 = d1;
    T.d2 = d2;
    edge_test(T.xy, v, e, 0, 1, T.n[2] < 0.0f, info.unit[0], info.unit[1]);
    edge_test(T.yz, v, e, 1, 2, T.n[0] < 0.0f, info.unit[1], info.unit[2]);
    edge_test(T.zx, v, e, 2, 0, T.n[1] < 0.0f, info.unit[0], info.unit[2]);
}

__device__ __forceinline__ void surface_setup(const float v[3][3], const vox_info& info, SurfTri& T) {
    surface_box(v, info, T.lo, T.hi);
    surface_setup_tests(v, info, T);
}

// Projection pass: all three edges of one projection accept point (pu, pv).
__device__ __forceinline__ bool edges_pass(const EdgeTest& E, float pu, float pv) {
    bool ok = true;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        float t = E.a[i] * pu;
        t = t + E.b[i] * pv;
        t = t + E.d[i];
        ok = ok && !(t < 0.0f);
    }
    return ok;
}

// Plane test for the voxel whose min corner is p: not rejected when
// (n.p + d1) * (n.p + d2) > 0 is false (NaN normals never reject).
__device__ __forceinline__ bool plane_pass(const SurfTri& T, float px, float py, float pz) {
    float t = T.n[0] * px;
    t = t + T.n[1] * py;
    t = t + T.n[2] * pz;
    float r = (t + T.d1) * (t + T.d2);
    return !(r > 0.0f);
}

__device__ __forceinline__ bool voxel_pass(const SurfTri& T, float px, float py, float pz) {
    return plane_pass(T, px, py, pz) && edges_pass(T.xy, px, py) && edges_pass(T.yz, py, pz) &&
           edges_pass(T.zx, pz, px);
}

// ---- solid (cpu_voxelizer.cpp:196-310) ----------------------------------------
struct SolidTri {
    float n[3];       // unit normal
    float v0[3];      // translated vertex 0 (x used by the x-intercept)
    float A[2], B[2], C[2];  // YZ projection, counter-clockwise
    float ymin, ymax, zmin, zmax;  // integer-valued column bounds (floats)
    uint32_t tl;      // top-left rule of edges AB, BC, CA (bits 0, 1, 2)
    bool valid;       // false when |n.x| < float_error (triangle parallel to x)
};

// check_point_triangle + top-left rule (cpu_voxelizer.cpp:196-238, :294)
__device__ __forceinline__ bool top_left(float p0, float p1, float q0, float q1) {
    return (q1 < p1) || (q1 == p1 && p0 > q0);
}

__device__ __forceinline__ void solid_setup(const float v[3][3], const vox_info& info, SolidTri& S) {
    float e0[3], e1[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        e0[k] = v[1][k] - v[0][k];
        e1[k] = v[2][k] - v[1][k];
    }
    unit_normal(e0, e1, S.n);
    S.valid = !((double)fabsf(S.n[0]) < kFloatError);
#pragma unroll
    for (int k = 0; k < 3; k++) S.v0[k] = v[0][k];
    float A0 = v[0][1], A1 = v[0][2];
    float B0 = v[1][1], B1 = v[1][2];
    float C0 = v[2][1], C1 = v[2][2];
    // checkCCW(A, B, C): (B-A).x*(C-A).y - (C-A).x*(B-A).y > 0
    float bx = B0 - A0, by = B1 - A1, cx = C0 - A0, cy = C1 - A1;
    float r = bx * cy - cx * by;
    if (!(r > 0.0f)) {
        float t0 = B0, t1 = B1;
        B0 = C0; B1 = C1;
        C0 = t0; C1 = t1;
    }
    S.A[0] = A0; S.A[1] = A1;
    S.B[0] = B0; S.B[1] = B1;
    S.C[0] = C0; S.C[1] = C1;
    float mxy = hmax(A0, hmax(B0, C0)), mxz = hmax(A1, hmax(B1, C1));
    float mny = hmin(A0, hmin(B0, C0)), mnz = hmin(A1, hmin(B1, C1));
    S.ymax = floorf(mxy / info.unit[1] - 0.5f);
    S.zmax = floorf(mxz / info.unit[2] - 0.5f);
    S.ymin = ceilf(mny / info.unit[1] - 0.5f);
    S.zmin = ceilf(mnz / info.unit[2] - 0.5f);
    S.tl = (top_left(A0, A1, B0, B1) ? 1u : 0u) | (top_left(B0, B1, C0, C1) ? 2u : 0u) |
           (top_left(C0, C1, A0, A1) ? 4u : 0u);
}

// (double)f < 1e-6 for a float f is f < 0x1.0c6f7cp-20f, the smallest float
// above 1e-6 (0x1.0c6f7ap-20f, the float below, is less than 1e-6).
__device__ __forceinline__ bool below_float_error(float f) { return fabsf(f) < 0x1.0c6f7cp-20f; }

// The reference tests the three edges in order and returns at the first on
// which the point lies (cpu_voxelizer.cpp:196-238); here every test is
// evaluated (no side effects, same operations) and the first one selected,
// which keeps the column loops free of divergent branches.
__device__ __forceinline__ bool column_hit(const SolidTri& S, float py, float pz) {
    const float PAx = py - S.A[0], PAy = pz - S.A[1];
    const float PBx = py - S.B[0], PBy = pz - S.B[1];
    const float PCx = py - S.C[0], PCy = pz - S.C[1];
    const float t1 = PAx * PBy - PAy * PBx;
    const float t2 = PBx * PCy - PBy * PCx;
    const float t3 = PCx * PAy - PCy * PAx;
    const bool on1 = below_float_error(t1) & (PAx * PBx <= 0.0f) & (PAy * PBy <= 0.0f);
    const bool on2 = below_float_error(t2) & (PBx * PCx <= 0.0f) & (PBy * PCy <= 0.0f);
    const bool on3 = below_float_error(t3) & (PCx * PAx <= 0.0f) & (PCy * PAy <= 0.0f);
    const bool inside = (t1 * t2 > 0.0f) & (t1 * t3 > 0.0f);
    const uint32_t edge = on1 ? 1u : (on2 ? 2u : (on3 ? 4u : 0u));
    return edge ? (S.tl & edge) != 0u : inside;
}

// x-run end of a hit column: int(float(xc / unit.x) - 0.5) with the
// subtraction in double (cpu_voxelizer.cpp:296).  Returns -1 for an empty run
// (q <= -1 or NaN: the reference CPU path would wrap its unsigned xmax; its GPU
// path, voxelize_solid.cu:123, writes nothing) and otherwise the run end, which
// may be >= G (the run then continues along the linear index into the next
// rows, as the reference indexes it).
__device__ __forceinline__ int64_t column_xmax(const SolidTri& S, const vox_info& info, float py, float pz) {
    float a = S.n[1] * (py - S.v0[1]);
    float b = S.n[2] * (pz - S.v0[2]);
    float s = a + b;
    float xc = (-s) / S.n[0] + S.v0[0];
    const double q = (double)(xc / info.unit[0]) - 0.5;
    const double qc = q < 4.0e18 ? q : 4.0e18;
    return q > -1.0 ? (int64_t)qc : (int64_t)-1;
}

// ---- index layouts (Appendix: MSB-first bits, word = idx >> 5) ---------------
// Interleave the low 16 bits of x, y, z: x_i -> bit 3i, y_i -> 3i+1, z_i -> 3i+2.
// Equal to the reference LUT encoder (voxelize.cuh:20-34), whose byte-2 term is
// shifted out, for every coordinate (both keep exactly the low 16 bits).
__device__ __forceinline__ uint64_t spread3(uint32_t v) {
    uint64_t x = v & 0xFFFFu;
    x = (x | (x << 16)) & 0x0000FF0000FFull;
    x = (x | (x << 8)) & 0x00F00F00F00Full;
    x = (x | (x << 4)) & 0x0C30C30C30C3ull;
    x = (x | (x << 2)) & 0x249249249249ull;
    return x;
}
__device__ __forceinline__ uint64_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
}
__device__ __forceinline__ uint64_t voxel_index(int layout, uint32_t G, int x, int y, int z) {
    if (layout == VOX_LAYOUT_MORTON) return morton3((uint32_t)x, (uint32_t)y, (uint32_t)z);
    return (uint64_t)x + (uint64_t)y * G + (uint64_t)z * G * G;
}
__device__ __forceinline__ uint32_t bit_mask(uint64_t idx) { return 1u << (31u - (uint32_t)(idx & 31u)); }

// XOR of the linear bit range [b0, b1] (word masks, atomics).
__device__ __forceinline__ void xor_linear_run(uint32_t* table, uint64_t b0, uint64_t b1) {
    for (uint64_t w = b0 >> 5; w <= (b1 >> 5); w++) {
        const uint32_t s = (w == (b0 >> 5)) ? (uint32_t)(b0 & 31u) : 0u;
        const uint32_t e = (w == (b1 >> 5)) ? (uint32_t)(b1 & 31u) : 31u;
        atomicXor(table + w, (0xFFFFFFFFu >> s) & (0xFFFFFFFFu << (31u - e)));
    }
}

// Full XOR run of one hit column (wide triangles), reference indexing.
__device__ __forceinline__ void xor_column_run(uint32_t* table, int layout, int G, int y, int z, int64_t xmax) {
    if (layout == VOX_LAYOUT_LINEAR) {
        const uint64_t total = (uint64_t)G * (uint64_t)G * (uint64_t)G;
        const uint64_t b0 = ((uint64_t)y + (uint64_t)z * (uint64_t)G) * (uint64_t)G;
        const uint64_t room = total - 1 - b0;
        xor_linear_run(table, b0, b0 + ((uint64_t)xmax < room ? (uint64_t)xmax : room));
    } else {
        const int xe = xmax < G ? (int)xmax : G - 1;
        uint64_t cur = ~0ull;
        uint32_t mask = 0;
        for (int x = 0; x <= xe; x++) {
            const uint64_t idx = morton3((uint32_t)x, (uint32_t)y, (uint32_t)z);
            if ((idx >> 5) != cur) {
                if (mask) atomicXor(table + cur, mask);
                cur = idx >> 5;
                mask = 0;
            }
            mask |= bit_mask(idx);
        }
        if (mask) atomicXor(table + cur, mask);
    }
}

}  // namespace voxamd
