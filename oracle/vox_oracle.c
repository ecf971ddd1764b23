/*
 * vox_oracle.c -- CPU restatement of the reference mesh voxelizer.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: tests/, the
 * smoke() check and bench.py's cpu_baseline leg may load it, nothing in the
 * product (cuda_voxelizer_amd/) may.  It restates, in plain C, the algorithm of
 * the reference's CPU voxelizer (Forceflow/cuda_voxelizer v0.6):
 *
 *   src/cpu_voxelizer.cpp:35-183   cpu_voxelize_mesh        (surface)
 *   src/cpu_voxelizer.cpp:241-314  cpu_voxelize_mesh_solid  (solid / parity)
 *   src/cpu_voxelizer.cpp:7-15     setBit     (MSB-first, omp critical)
 *   src/cpu_voxelizer.cpp:186-194  setBitXor
 *   src/cpu_voxelizer.cpp:18-32    mortonEncode_LUT (host)
 *   src/cpu_voxelizer.cpp:196-238  TopLeftEdge / checkCCW / get_x_coordinate /
 *                                  check_point_triangle
 *   src/util.h:20-22,50-61,80-110  float3_to_int3, voxinfo, createMeshBBCube
 *   src/libs/cuda/helper_math.h:51-81 host fallbacks (fminf/fmaxf ternaries,
 *                                  rsqrtf = 1/sqrtf) and the float2/float3 ops
 *   src/timer.h:13                 `using namespace std` -> max(0.0f,a) is
 *                                  std::max<float>, i.e. (0.0f < a) ? a : 0.0f
 *
 * Every floating-point expression is written in the reference's textual
 * order, one binary32 operation at a time; the file must be compiled with
 * -ffp-contract=off and without -ffast-math / -ffinite-math-only (NaN normals
 * of zero-area triangles are part of the semantics).
 *
 * Two execution modes give identical tables:
 *   VOXO_MODE_REF  : omp parallel for over triangles, one global critical
 *                    section per bit write (the reference's structure; this is
 *                    the timed CPU baseline).
 *   VOXO_MODE_FAST : same loops, lock-free __atomic OR/XOR per bit (the
 *                    expected tables of the parity tests at large sizes).
 *
 * Out-of-row runs (documented, counted).  The reference's point-in-triangle
 * test accepts column centres within an absolute 1e-6 (world units^2) of an
 * edge, so at fine grids a column slightly outside a triangle that is nearly
 * parallel to x can be accepted and its extrapolated x-intercept can leave the
 * row.  The reference CPU path then stores xmax in an unsigned int:
 *   xmax >= G : the run x = 0..xmax continues through the following linear
 *               indices (x + y*G + z*G*G), i.e. into the next rows -- kept
 *               here bit for bit; only bits past the end of the table (which
 *               the reference would write out of bounds) are dropped.  In the
 *               Morton layout every x >= G maps past the table, so the run
 *               stops at G-1.
 *   xmax < 0  : the unsigned wrap makes the reference loop ~2^32 times and
 *               write out of bounds (a crash); here, as in the reference GPU
 *               path (voxelize_solid.cu:123, int xmax), the run is empty.
 * Every such run, and any column index outside the grid (never produced for a
 * mesh inside the voxinfo box), is counted in *out_offrow.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <omp.h>

#define VOXO_MODE_REF 0
#define VOXO_MODE_FAST 1

/* reference: #define float_error 0.000001  (a double literal), cpu_voxelizer.cpp:2 */
static const double float_error = 0.000001;

/* Layout mirror of the reference voxinfo (util.h:50-61): 64 bytes.
   0: bbox.min (3 f32), 12: bbox.max, 24: gridsize (3 u32), 40: n_triangles (u64),
   48: unit (3 f32). */
typedef struct {
    float bmin[3];
    float bmax[3];
    uint32_t grid[3];
    uint32_t _pad0;
    uint64_t n_triangles;
    float unit[3];
    uint32_t _pad1;
} voxo_info;

typedef struct { float x, y, z; } f3;
typedef struct { float x, y; } f2;
typedef struct { int x, y, z; } i3;

/* ---- helper_math.h host semantics (restated) ---------------------------- */
static inline f3 mk3(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
static inline f2 mk2(float x, float y) { f2 r; r.x = x; r.y = y; return r; }
static inline f3 sub3(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline f2 sub2(f2 a, f2 b) { return mk2(a.x - b.x, a.y - b.y); }
static inline f3 div3(f3 a, f3 b) { return mk3(a.x / b.x, a.y / b.y, a.z / b.z); }
static inline f3 muls3(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
/* dot: a.x*b.x + a.y*b.y + a.z*b.z, left-associative (helper_math.h:1264-1267) */
static inline float dot3(f3 a, f3 b) { float t = a.x * b.x; t = t + a.y * b.y; t = t + a.z * b.z; return t; }
static inline float dot2(f2 a, f2 b) { float t = a.x * b.x; t = t + a.y * b.y; return t; }
/* cross (helper_math.h:1436-1439) */
static inline f3 cross3(f3 a, f3 b) {
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
/* host rsqrtf = 1.0f / sqrtf(x) (helper_math.h:78-81); normalize = v * rsqrtf(dot(v,v)) */
static inline f3 normalize3(f3 v) { float inv = 1.0f / sqrtf(dot3(v, v)); return muls3(v, inv); }
/* host fminf / fmaxf are ternaries, not IEEE minNum (helper_math.h:58-66) */
static inline float hfminf(float a, float b) { return a < b ? a : b; }
static inline float hfmaxf(float a, float b) { return a > b ? a : b; }
static inline f3 fmin3(f3 a, f3 b) { return mk3(hfminf(a.x, b.x), hfminf(a.y, b.y), hfminf(a.z, b.z)); }
static inline f3 fmax3(f3 a, f3 b) { return mk3(hfmaxf(a.x, b.x), hfmaxf(a.y, b.y), hfmaxf(a.z, b.z)); }
static inline f2 fmin2(f2 a, f2 b) { return mk2(hfminf(a.x, b.x), hfminf(a.y, b.y)); }
static inline f2 fmax2(f2 a, f2 b) { return mk2(hfmaxf(a.x, b.x), hfmaxf(a.y, b.y)); }
/* std::max<float>(0.0f, a) == (0.0f < a) ? a : 0.0f */
static inline float stdmax0(float a) { return (0.0f < a) ? a : 0.0f; }
/* int clamp: max(a, min(f, b)) (helper_math.h:1218-1221) */
static inline int iclamp(int f, int a, int b) { int m = f < b ? f : b; return a > m ? a : m; }
/* util.h:20-22 float3_to_int3: static_cast<int> truncation */
static inline i3 f3_to_i3(f3 a) { i3 r; r.x = (int)a.x; r.y = (int)a.y; r.z = (int)a.z; return r; }

/* ---- Morton (cpu_voxelizer.cpp:18-32, morton_LUTs.h) --------------------- */
/* The reference ships 3x256 LUTs; entry i of the x table spreads bit k of i to
   bit 3k, the y table is x<<1 and the z table x<<2 (morton_LUTs.h:5,41,77).
   They are regenerated here from that closed form. */
static uint32_t lut_x[256], lut_y[256], lut_z[256];
static int lut_ready = 0;

static void build_luts(void) {
    if (lut_ready) return;
    for (uint32_t i = 0; i < 256; i++) {
        uint32_t v = 0;
        for (int k = 0; k < 8; k++) v |= ((i >> k) & 1u) << (3 * k);
        lut_x[i] = v;
        lut_y[i] = v << 1;
        lut_z[i] = v << 2;
    }
    lut_ready = 1;
}

/* Literal restatement: three byte terms combined with <<48 and <<24 on a
   64-bit accumulator; the byte-2 term is shifted out entirely, so the code is
   the true interleave only for coordinates < 2^16. */
uint64_t voxo_morton_encode(uint32_t x, uint32_t y, uint32_t z) {
    uint64_t answer = 0;
    build_luts();
    answer = (uint64_t)(lut_z[(z >> 16) & 0xFF] | lut_y[(y >> 16) & 0xFF] | lut_x[(x >> 16) & 0xFF]);
    answer = answer << 48 | lut_z[(z >> 8) & 0xFF] | lut_y[(y >> 8) & 0xFF] | lut_x[(x >> 8) & 0xFF];
    answer = answer << 24 | lut_z[z & 0xFF] | lut_y[y & 0xFF] | lut_x[x & 0xFF];
    return answer;
}

void voxo_morton_luts(uint32_t* out768) {
    build_luts();
    memcpy(out768, lut_x, sizeof lut_x);
    memcpy(out768 + 256, lut_y, sizeof lut_y);
    memcpy(out768 + 512, lut_z, sizeof lut_z);
}

/* ---- bit writes (cpu_voxelizer.cpp:7-15, 186-194) ------------------------ */
static inline void set_bit(uint32_t* table, size_t index, int mode) {
    size_t int_location = index / (size_t)32;
    uint32_t bit_pos = (uint32_t)((size_t)31 - (index % (size_t)32));
    uint32_t mask = 1u << bit_pos;
    if (mode == VOXO_MODE_REF) {
#pragma omp critical
        { table[int_location] = (table[int_location] | mask); }
    } else {
        __atomic_fetch_or(&table[int_location], mask, __ATOMIC_RELAXED);
    }
}

static inline void set_bit_xor(uint32_t* table, size_t index, int mode) {
    size_t int_location = index / (size_t)32;
    uint32_t bit_pos = (uint32_t)((size_t)31 - (index % (size_t)32));
    uint32_t mask = 1u << bit_pos;
    if (mode == VOXO_MODE_REF) {
#pragma omp critical
        { table[int_location] = (table[int_location] ^ mask); }
    } else {
        __atomic_fetch_xor(&table[int_location], mask, __ATOMIC_RELAXED);
    }
}

static inline size_t linear_index(const voxo_info* info, int x, int y, int z) {
    /* cpu_voxelizer.cpp:168 / :304 (x + y*G.y + z*G.y*G.z; equal to the
       x-major index because grids are cubic) */
    return (size_t)x + ((size_t)y * (size_t)info->grid[1]) +
           ((size_t)z * (size_t)info->grid[1] * (size_t)info->grid[2]);
}

/* REF mode keeps the reference's default (static) schedule; FAST mode balances
   uneven triangle sizes dynamically.  The schedule never changes the table. */
static void set_schedule(int mode) {
    if (mode == VOXO_MODE_REF) omp_set_schedule(omp_sched_static, 0);
    else omp_set_schedule(omp_sched_dynamic, 256);
}

static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec * 1e3 + (double)ts.tv_nsec * 1e-6;
}

/* Translate prepass, cpu_voxelizer.cpp:40-45 / 246-251: vertices -= bbox.min
   (in place, as the reference mutates its TriMesh). */
static void translate_prepass(const voxo_info* info, float* verts, int64_t nverts) {
#pragma omp parallel for
    for (int64_t i = 0; i < nverts; i++) {
        verts[3 * i + 0] = verts[3 * i + 0] - info->bmin[0];
        verts[3 * i + 1] = verts[3 * i + 1] - info->bmin[1];
        verts[3 * i + 2] = verts[3 * i + 2] - info->bmin[2];
    }
}

static inline f3 vtx(const float* verts, int32_t idx) {
    return mk3(verts[3 * (int64_t)idx + 0], verts[3 * (int64_t)idx + 1], verts[3 * (int64_t)idx + 2]);
}

/* ---- surface: cpu_voxelizer.cpp:35-183 ----------------------------------- */
/* verts: float[3*nverts] (modified in place by the prepass); faces: int32[3*n]
   with n = info->n_triangles.  Returns the elapsed milliseconds of prepass +
   triangle loop (the span the reference's Timer covers). */
double voxo_surface(const voxo_info* info, float* verts, int64_t nverts, const int32_t* faces,
                    uint32_t* table, int morton, int mode) {
    build_luts();
    double t0 = now_ms();
    translate_prepass(info, verts, nverts);
    const int64_t ntri = (int64_t)info->n_triangles;
    set_schedule(mode);

#pragma omp parallel for schedule(runtime)
    for (int64_t i = 0; i < ntri; i++) {
        const f3 unit = mk3(info->unit[0], info->unit[1], info->unit[2]);
        f3 delta_p = unit;
        f3 c = mk3(0.0f, 0.0f, 0.0f);
        i3 grid_max;
        grid_max.x = (int)info->grid[0] - 1;
        grid_max.y = (int)info->grid[1] - 1;
        grid_max.z = (int)info->grid[2] - 1;

        f3 v0 = vtx(verts, faces[3 * i + 0]);
        f3 v1 = vtx(verts, faces[3 * i + 1]);
        f3 v2 = vtx(verts, faces[3 * i + 2]);

        f3 e0 = sub3(v1, v0);
        f3 e1 = sub3(v2, v1);
        f3 e2 = sub3(v0, v2);
        f3 n = normalize3(cross3(e0, e1));

        /* :76-80 triangle bbox in world, then grid (truncate, clamp) */
        f3 wmin = fmin3(v0, fmin3(v1, v2));
        f3 wmax = fmax3(v0, fmax3(v1, v2));
        i3 gmin = f3_to_i3(div3(wmin, unit));
        i3 gmax = f3_to_i3(div3(wmax, unit));
        gmin.x = iclamp(gmin.x, 0, grid_max.x);
        gmin.y = iclamp(gmin.y, 0, grid_max.y);
        gmin.z = iclamp(gmin.z, 0, grid_max.z);
        gmax.x = iclamp(gmax.x, 0, grid_max.x);
        gmax.y = iclamp(gmax.y, 0, grid_max.y);
        gmax.z = iclamp(gmax.z, 0, grid_max.z);

        /* :83-87 plane test constants (critical point c) */
        if (n.x > 0.0f) { c.x = unit.x; }
        if (n.y > 0.0f) { c.y = unit.y; }
        if (n.z > 0.0f) { c.z = unit.z; }
        float d1 = dot3(n, sub3(c, v0));
        float d2 = dot3(n, sub3(sub3(delta_p, c), v0));

        /* :91-125 projection edge normals and offsets */
        f2 n_xy_e0 = mk2(-1.0f * e0.y, e0.x);
        f2 n_xy_e1 = mk2(-1.0f * e1.y, e1.x);
        f2 n_xy_e2 = mk2(-1.0f * e2.y, e2.x);
        if (n.z < 0.0f) {
            n_xy_e0 = mk2(-n_xy_e0.x, -n_xy_e0.y);
            n_xy_e1 = mk2(-n_xy_e1.x, -n_xy_e1.y);
            n_xy_e2 = mk2(-n_xy_e2.x, -n_xy_e2.y);
        }
        float d_xy_e0 = (-1.0f * dot2(n_xy_e0, mk2(v0.x, v0.y))) + stdmax0(unit.x * n_xy_e0.x) + stdmax0(unit.y * n_xy_e0.y);
        float d_xy_e1 = (-1.0f * dot2(n_xy_e1, mk2(v1.x, v1.y))) + stdmax0(unit.x * n_xy_e1.x) + stdmax0(unit.y * n_xy_e1.y);
        float d_xy_e2 = (-1.0f * dot2(n_xy_e2, mk2(v2.x, v2.y))) + stdmax0(unit.x * n_xy_e2.x) + stdmax0(unit.y * n_xy_e2.y);

        f2 n_yz_e0 = mk2(-1.0f * e0.z, e0.y);
        f2 n_yz_e1 = mk2(-1.0f * e1.z, e1.y);
        f2 n_yz_e2 = mk2(-1.0f * e2.z, e2.y);
        if (n.x < 0.0f) {
            n_yz_e0 = mk2(-n_yz_e0.x, -n_yz_e0.y);
            n_yz_e1 = mk2(-n_yz_e1.x, -n_yz_e1.y);
            n_yz_e2 = mk2(-n_yz_e2.x, -n_yz_e2.y);
        }
        float d_yz_e0 = (-1.0f * dot2(n_yz_e0, mk2(v0.y, v0.z))) + stdmax0(unit.y * n_yz_e0.x) + stdmax0(unit.z * n_yz_e0.y);
        float d_yz_e1 = (-1.0f * dot2(n_yz_e1, mk2(v1.y, v1.z))) + stdmax0(unit.y * n_yz_e1.x) + stdmax0(unit.z * n_yz_e1.y);
        float d_yz_e2 = (-1.0f * dot2(n_yz_e2, mk2(v2.y, v2.z))) + stdmax0(unit.y * n_yz_e2.x) + stdmax0(unit.z * n_yz_e2.y);

        f2 n_zx_e0 = mk2(-1.0f * e0.x, e0.z);
        f2 n_zx_e1 = mk2(-1.0f * e1.x, e1.z);
        f2 n_zx_e2 = mk2(-1.0f * e2.x, e2.z);
        if (n.y < 0.0f) {
            n_zx_e0 = mk2(-n_zx_e0.x, -n_zx_e0.y);
            n_zx_e1 = mk2(-n_zx_e1.x, -n_zx_e1.y);
            n_zx_e2 = mk2(-n_zx_e2.x, -n_zx_e2.y);
        }
        /* unit axes exactly as the reference writes them (:123-125) */
        float d_xz_e0 = (-1.0f * dot2(n_zx_e0, mk2(v0.z, v0.x))) + stdmax0(unit.x * n_zx_e0.x) + stdmax0(unit.z * n_zx_e0.y);
        float d_xz_e1 = (-1.0f * dot2(n_zx_e1, mk2(v1.z, v1.x))) + stdmax0(unit.x * n_zx_e1.x) + stdmax0(unit.z * n_zx_e1.y);
        float d_xz_e2 = (-1.0f * dot2(n_zx_e2, mk2(v2.z, v2.x))) + stdmax0(unit.x * n_zx_e2.x) + stdmax0(unit.z * n_zx_e2.y);

        /* :128-175 voxel loop, z -> y -> x */
        for (int z = gmin.z; z <= gmax.z; z++) {
            for (int y = gmin.y; y <= gmax.y; y++) {
                for (int x = gmin.x; x <= gmax.x; x++) {
                    f3 p = mk3((float)x * unit.x, (float)y * unit.y, (float)z * unit.z);
                    float nDOTp = dot3(n, p);
                    if (((nDOTp + d1) * (nDOTp + d2)) > 0.0f) { continue; }

                    f2 p_xy = mk2(p.x, p.y);
                    if ((dot2(n_xy_e0, p_xy) + d_xy_e0) < 0.0f) { continue; }
                    if ((dot2(n_xy_e1, p_xy) + d_xy_e1) < 0.0f) { continue; }
                    if ((dot2(n_xy_e2, p_xy) + d_xy_e2) < 0.0f) { continue; }

                    f2 p_yz = mk2(p.y, p.z);
                    if ((dot2(n_yz_e0, p_yz) + d_yz_e0) < 0.0f) { continue; }
                    if ((dot2(n_yz_e1, p_yz) + d_yz_e1) < 0.0f) { continue; }
                    if ((dot2(n_yz_e2, p_yz) + d_yz_e2) < 0.0f) { continue; }

                    f2 p_zx = mk2(p.z, p.x);
                    if ((dot2(n_zx_e0, p_zx) + d_xz_e0) < 0.0f) { continue; }
                    if ((dot2(n_zx_e1, p_zx) + d_xz_e1) < 0.0f) { continue; }
                    if ((dot2(n_zx_e2, p_zx) + d_xz_e2) < 0.0f) { continue; }

                    if (morton) {
                        set_bit(table, (size_t)voxo_morton_encode((uint32_t)x, (uint32_t)y, (uint32_t)z), mode);
                    } else {
                        set_bit(table, linear_index(info, x, y, z), mode);
                    }
                }
            }
        }
    }
    return now_ms() - t0;
}

/* ---- solid helpers: cpu_voxelizer.cpp:196-238 ----------------------------- */
static inline int top_left_edge(f2 v0, f2 v1) {
    return ((v1.y < v0.y) || (v1.y == v0.y && v0.x > v1.x));
}

static inline int check_ccw(f2 v0, f2 v1, f2 v2) {
    f2 e0 = sub2(v1, v0);
    f2 e1 = sub2(v2, v0);
    float result = e0.x * e1.y - e1.x * e0.y;
    return result > 0;
}

static inline float get_x_coordinate(f3 n, f3 v0, f2 point) {
    /* (-(n.y * (point.x - v0.y) + n.z * (point.y - v0.z)) / n.x + v0.x) */
    float a = n.y * (point.x - v0.y);
    float b = n.z * (point.y - v0.z);
    float s = a + b;
    return (-s) / n.x + v0.x;
}

static inline int check_point_triangle(f2 v0, f2 v1, f2 v2, f2 point) {
    f2 PA = sub2(point, v0);
    f2 PB = sub2(point, v1);
    f2 PC = sub2(point, v2);

    float t1 = PA.x * PB.y - PA.y * PB.x;
    if ((double)fabsf(t1) < float_error && PA.x * PB.x <= 0 && PA.y * PB.y <= 0) return 1;

    float t2 = PB.x * PC.y - PB.y * PC.x;
    if ((double)fabsf(t2) < float_error && PB.x * PC.x <= 0 && PB.y * PC.y <= 0) return 2;

    float t3 = PC.x * PA.y - PC.y * PA.x;
    if ((double)fabsf(t3) < float_error && PC.x * PA.x <= 0 && PC.y * PA.y <= 0) return 3;

    if (t1 * t2 > 0 && t1 * t3 > 0) return 0;
    return -1;
}

/* ---- solid: cpu_voxelizer.cpp:241-314 ------------------------------------- */
double voxo_solid(const voxo_info* info, float* verts, int64_t nverts, const int32_t* faces,
                  uint32_t* table, int morton, int mode, int64_t* out_offrow) {
    build_luts();
    double t0 = now_ms();
    translate_prepass(info, verts, nverts);
    const int64_t ntri = (int64_t)info->n_triangles;
    const int G0 = (int)info->grid[0], G1 = (int)info->grid[1], G2 = (int)info->grid[2];
    const size_t total_bits = (size_t)G0 * (size_t)G1 * (size_t)G2;
    int64_t offrow = 0;
    set_schedule(mode);

#pragma omp parallel for schedule(runtime) reduction(+ : offrow)
    for (int64_t i = 0; i < ntri; i++) {
        const f3 unit = mk3(info->unit[0], info->unit[1], info->unit[2]);
        f3 v0 = vtx(verts, faces[3 * i + 0]);
        f3 v1 = vtx(verts, faces[3 * i + 1]);
        f3 v2 = vtx(verts, faces[3 * i + 2]);
        f3 e0 = sub3(v1, v0);
        f3 e1 = sub3(v2, v1);
        f3 n = normalize3(cross3(e0, e1));
        /* std::fabs(float) promoted to double against 1e-6 (:265) */
        if ((double)fabsf(n.x) < float_error) { continue; }

        f2 v0_yz = mk2(v0.y, v0.z);
        f2 v1_yz = mk2(v1.y, v1.z);
        f2 v2_yz = mk2(v2.y, v2.z);
        if (!check_ccw(v0_yz, v1_yz, v2_yz)) {
            f2 v3 = v1_yz;
            v1_yz = v2_yz;
            v2_yz = v3;
        }

        f2 bbox_max = fmax2(v0_yz, fmax2(v1_yz, v2_yz));
        f2 bbox_min = fmin2(v0_yz, fmin2(v1_yz, v2_yz));
        /* :285-286 float arithmetic: q / unit - 0.5f, then floor/ceil */
        f2 bbox_max_grid = mk2(floorf(bbox_max.x / unit.y - 0.5f), floorf(bbox_max.y / unit.z - 0.5f));
        f2 bbox_min_grid = mk2(ceilf(bbox_min.x / unit.y - 0.5f), ceilf(bbox_min.y / unit.z - 0.5f));

        /* :288-290, with the documented clamp to the grid (counted) */
        int ylo = (int)bbox_min_grid.x, zlo = (int)bbox_min_grid.y;
        if (ylo < 0) { offrow++; ylo = 0; }
        if (zlo < 0) { offrow++; zlo = 0; }
        if (bbox_max_grid.x > (float)(G1 - 1)) { offrow++; }
        if (bbox_max_grid.y > (float)(G2 - 1)) { offrow++; }
        for (int y = ylo; (float)y <= bbox_max_grid.x && y < G1; y++) {
            for (int z = zlo; (float)z <= bbox_max_grid.y && z < G2; z++) {
                f2 point = mk2(((float)y + 0.5f) * unit.y, ((float)z + 0.5f) * unit.z);
                int checknum = check_point_triangle(v0_yz, v1_yz, v2_yz, point);
                if ((checknum == 1 && top_left_edge(v0_yz, v1_yz)) ||
                    (checknum == 2 && top_left_edge(v1_yz, v2_yz)) ||
                    (checknum == 3 && top_left_edge(v2_yz, v0_yz)) || (checknum == 0)) {
                    /* :296 int(float / float - 0.5) -- the subtraction is in double */
                    double q = (double)(get_x_coordinate(n, v0, point) / unit.x) - 0.5;
                    if (!(q > -1.0)) { offrow++; continue; }  /* negative or NaN: empty run */
                    int64_t xmax = (q < 4.0e18) ? (int64_t)q : (int64_t)4.0e18;
                    if (xmax >= G0) offrow++;
                    if (morton) {
                        if (xmax > G0 - 1) xmax = G0 - 1;  /* x >= G is past the table */
                        for (int64_t x = 0; x <= xmax; x++)
                            set_bit_xor(table, (size_t)voxo_morton_encode((uint32_t)x, (uint32_t)y, (uint32_t)z), mode);
                    } else {
                        const size_t base = linear_index(info, 0, y, z);
                        for (int64_t x = 0; x <= xmax && base + (size_t)x < total_bits; x++)
                            set_bit_xor(table, base + (size_t)x, mode);
                    }
                }
            }
        }
    }
    if (out_offrow) *out_offrow = offrow;
    return now_ms() - t0;
}

/* ---- host prep: util.h:50-61 (voxinfo) and util.h:80-110 (createMeshBBCube) */
/* Component-wise min/max over all vertices (trimesh2 need_bbox semantics, the
   third-party call at main.cpp:179). */
void voxo_mesh_bbox(const float* verts, int64_t nverts, float* out_min3, float* out_max3) {
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = 0; i < nverts; i++) {
        for (int k = 0; k < 3; k++) {
            float v = verts[3 * i + k];
            if (v < mn[k]) mn[k] = v;
            if (v > mx[k]) mx[k] = v;
        }
    }
    memcpy(out_min3, mn, sizeof mn);
    memcpy(out_max3, mx, sizeof mx);
}

void voxo_bbox_cube(const float* in_min3, const float* in_max3, float* out_min3, float* out_max3) {
    float amin[3] = {in_min3[0], in_min3[1], in_min3[2]};
    float amax[3] = {in_max3[0], in_max3[1], in_max3[2]};
    float len[3] = {in_max3[0] - in_min3[0], in_max3[1] - in_min3[1], in_max3[2] - in_min3[2]};
    /* std::max(lengths.x, std::max(lengths.y, lengths.z)); std::max(a,b) = (a<b)?b:a */
    float myz = (len[1] < len[2]) ? len[2] : len[1];
    float max_length = (len[0] < myz) ? myz : len[0];
    for (int k = 0; k < 3; k++) {
        if (max_length != len[k]) {
            float delta = max_length - len[k];
            amin[k] = in_min3[k] - (delta / 2.0f);
            amax[k] = in_max3[k] + (delta / 2.0f);
        }
    }
    for (int k = 0; k < 3; k++) {
        float eps = (amax[k] - amin[k]) / 10001.0f;
        out_min3[k] = amin[k] - eps;
        out_max3[k] = amax[k] + eps;
    }
}

void voxo_make_info(const float* bmin3, const float* bmax3, uint32_t grid, uint64_t n_triangles, voxo_info* out) {
    memset(out, 0, sizeof *out);
    for (int k = 0; k < 3; k++) {
        out->bmin[k] = bmin3[k];
        out->bmax[k] = bmax3[k];
        out->grid[k] = grid;
        out->unit[k] = (bmax3[k] - bmin3[k]) / (float)grid;
    }
    out->n_triangles = n_triangles;
}

int voxo_sizeof_info(void) { return (int)sizeof(voxo_info); }

void voxo_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }
int voxo_max_threads(void) { return omp_get_max_threads(); }
