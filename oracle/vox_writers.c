/*
 * vox_writers.c -- CPU restatement of the reference's output writers.
 *
 * TEST INFRASTRUCTURE ONLY (the parity oracle for the device encoders; the
 * product package never loads it).  Restates, with the same loops and the
 * same per-voxel checkVoxel reads, the file bodies that the reference's
 * writers produce (Forceflow/cuda_voxelizer v0.6):
 *
 *   src/util_io.cpp:202-246  write_binvox          header + (value, count) RLE
 *   src/util_io.cpp:154-190  write_obj_pointcloud  "v x+0.5 y+0.5 z+0.5"
 *   src/util_io.cpp:45-90    write_cube            8 "v" + 12 "f" lines
 *   src/util_io.cpp:92-152   write_obj_cubes       loop order (before its
 *                            trimesh2 reorder_verts re-write, not restated:
 *                            trimesh2 is not available here)
 *   src/util.h:25-38         checkVoxel            linear read-back
 *
 * std::ostream's default float formatting is printf's %g with precision 6
 * (the C++ num_put rules), so the text lines are produced with snprintf.
 * The reference's RLE counter is a `char` compared with (char)255; with the
 * usual wrap-around that caps runs at 255 voxels, which is what is written
 * here with an unsigned counter.  Each function returns the body length and
 * writes at most `cap` bytes (call with cap = 0 to size the output).
 */
#include <stdint.h>
#include <stdio.h>
#include <string.h>

uint64_t voxo_morton_encode(uint32_t x, uint32_t y, uint32_t z);

/* checkVoxel (util.h:25-38) for a cubic grid; Morton tables read the same bit
   at the Morton code (the layout the GPU and CPU voxelizers write). */
static int check_voxel(const uint32_t* t, uint64_t G, uint64_t x, uint64_t y, uint64_t z, int morton) {
    const uint64_t loc = morton ? voxo_morton_encode((uint32_t)x, (uint32_t)y, (uint32_t)z) : x + y * G + z * G * G;
    return (t[loc / 32] >> (31 - (loc % 32))) & 1u;
}

typedef struct {
    uint8_t* out;
    uint64_t cap, len;
} sink;

static void put(sink* s, const void* p, uint64_t n) {
    if (s->len + n <= s->cap) memcpy(s->out + s->len, p, n);
    s->len += n;
}

uint64_t voxo_binvox_body(const uint32_t* table, uint32_t grid, int morton, uint8_t* out, uint64_t cap) {
    sink s = {out, cap, 0};
    const uint64_t G = grid;
    uint8_t value = 0;
    unsigned seen = 0;
    for (uint64_t x = 0; x < G; x++) {
        for (uint64_t z = 0; z < G; z++) {
            for (uint64_t y = 0; y < G; y++) {
                if (x == 0 && y == 0 && z == 0) {
                    value = (uint8_t)check_voxel(table, G, 0, 0, 0, morton);
                    put(&s, &value, 1);
                    seen = 1;
                    continue;
                }
                const uint8_t next = (uint8_t)check_voxel(table, G, x, y, z, morton);
                if (next != value || seen == 255) {
                    const uint8_t c = (uint8_t)seen;
                    put(&s, &c, 1);
                    seen = 1;
                    value = next;
                    put(&s, &value, 1);
                } else {
                    seen++;
                }
            }
        }
    }
    const uint8_t c = (uint8_t)seen;
    put(&s, &c, 1);
    return s.len;
}

uint64_t voxo_obj_points(const uint32_t* table, uint32_t grid, int morton, uint8_t* out, uint64_t cap) {
    sink s = {out, cap, 0};
    const uint64_t G = grid;
    char line[128];
    for (uint64_t x = 0; x < G; x++)
        for (uint64_t y = 0; y < G; y++)
            for (uint64_t z = 0; z < G; z++)
                if (check_voxel(table, G, x, y, z, morton)) {
                    const int n = snprintf(line, sizeof line, "v %g %g %g\n", x + 0.5, y + 0.5, z + 0.5);
                    put(&s, line, (uint64_t)n);
                }
    return s.len;
}

static void cube(sink* s, int x, int y, int z) {
    /* vertices v1..v8 of write_cube, written v8 first */
    const int v[8][3] = {{x + 1, y + 1, z + 1}, {x, y + 1, z + 1}, {x + 1, y, z + 1}, {x, y, z + 1},
                         {x, y, z},             {x + 1, y, z},     {x, y + 1, z},     {x + 1, y + 1, z}};
    static const int f[12][3] = {{-1, -3, -4}, {-1, -4, -2}, {-4, -3, -6}, {-4, -6, -5}, {-3, -1, -8}, {-3, -8, -6},
                                 {-1, -2, -7}, {-1, -7, -8}, {-2, -4, -5}, {-2, -5, -7}, {-5, -6, -8}, {-5, -8, -7}};
    char line[96];
    for (int k = 7; k >= 0; k--) {
        const int n = snprintf(line, sizeof line, "v %d %d %d\n", v[k][0], v[k][1], v[k][2]);
        put(s, line, (uint64_t)n);
    }
    for (int k = 0; k < 12; k++) {
        const int n = snprintf(line, sizeof line, "f %d %d %d\n", f[k][0], f[k][1], f[k][2]);
        put(s, line, (uint64_t)n);
    }
}

uint64_t voxo_obj_cubes(const uint32_t* table, uint32_t grid, int morton, uint8_t* out, uint64_t cap) {
    sink s = {out, cap, 0};
    const uint64_t G = grid;
    for (uint64_t x = 0; x < G; x++)
        for (uint64_t y = 0; y < G; y++)
            for (uint64_t z = 0; z < G; z++)
                if (check_voxel(table, G, x, y, z, morton)) cube(&s, (int)x, (int)y, (int)z);
    return s.len;
}

/* ---- MagicaVoxel .vox (util_io.cpp:249-285 + libs/magicavoxel_file_writer) ----
 * write_vox feeds AddVoxel(x, G - z, y, 1) for every set voxel in x, y, z
 * order (z fastest) to a VoxWriter with 126^3 cubes (VoxWriter.h default
 * arguments), after AddColor(255, 255, 255, 0, 0).  Restated here:
 *   AddVoxel (VoxWriter.cpp:444-457): cube = floor(v / 126) per axis; the
 *     running minCubeX, and minCubeY/minCubeZ = min(minCubeX, oy/oz) of the
 *     latest voxel (the writer passes minCubeX, not minCubeY/Z);
 *     maxVolume grows to the voxel-coordinate bounding box.
 *   m_GetCube / m_MergeVoxelInCube (:653-716): cube ids in order of first
 *     use; voxel bytes (X % 126, Y % 126, Z % 126, colour) appended per cube.
 *   SaveToFile (:459-560): header, per cube a SIZE and an XYZI chunk, root
 *     nTRN, nGRP, per cube nTRN (frame "_t" = translation) + nSHP (model,
 *     "_f" = "0"), RGBA palette (entry 255 is left uninitialised by the
 *     writer; 0 here), MAIN child size patched at the end.
 * Float-to-int conversions out of range give INT_MIN (x86 truncation). */
#include <math.h>
#include <stdlib.h>

typedef struct {
    size_t ox, oy, oz;
    uint8_t* v;
    uint64_t n, cap;
} vcube;

static int32_t x86_int(double d) {
    if (!(d >= -2147483648.0 && d < 2147483648.0)) return INT32_MIN;
    return (int32_t)d;
}

static void put32(sink* s, int32_t v) {
    uint8_t b[4];
    for (int k = 0; k < 4; k++) b[k] = (uint8_t)((uint32_t)v >> (8 * k));
    put(s, b, 4);
}

static void putstr(sink* s, const char* str) {
    const int32_t n = (int32_t)strlen(str);
    put32(s, n);
    put(s, str, (uint64_t)n);
}

uint64_t voxo_vox_file(const uint32_t* table, uint32_t grid, int morton, uint8_t* out, uint64_t cap) {
    const size_t G = grid, M = 126;
    const size_t nc = G / M + 2;  /* cube coordinates per axis (Y = G - z reaches G) */
    int32_t* id = (int32_t*)malloc(sizeof(int32_t) * nc * nc * nc);
    for (size_t i = 0; i < nc * nc * nc; i++) id[i] = -1;
    vcube* cubes = NULL;
    size_t ncubes = 0, capc = 0;
    size_t minCubeX = (size_t)1e7, minCubeY = (size_t)1e7, minCubeZ = (size_t)1e7;
    double lo[3] = {1e7, 1e7, 1e7}, hi[3] = {-1e7, -1e7, -1e7};
    for (size_t x = 0; x < G; x++)
        for (size_t y = 0; y < G; y++)
            for (size_t z = 0; z < G; z++) {
                if (!check_voxel(table, G, x, y, z, morton)) continue;
                const size_t vX = x, vY = -z + G, vZ = y;
                const size_t ox = (size_t)floor((double)vX / (double)M);
                const size_t oy = (size_t)floor((double)vY / (double)M);
                const size_t oz = (size_t)floor((double)vZ / (double)M);
                minCubeX = minCubeX < ox ? minCubeX : ox;
                minCubeY = minCubeX < oy ? minCubeX : oy;
                minCubeZ = minCubeX < oz ? minCubeX : oz;
                const size_t key = (ox * nc + oy) * nc + oz;
                if (id[key] < 0) {
                    if (ncubes == capc) {
                        capc = capc ? 2 * capc : 16;
                        cubes = (vcube*)realloc(cubes, capc * sizeof(vcube));
                    }
                    vcube c = {ox, oy, oz, NULL, 0, 0};
                    cubes[ncubes] = c;
                    id[key] = (int32_t)ncubes++;
                }
                const double p[3] = {(double)vX, (double)vY, (double)vZ};
                for (int k = 0; k < 3; k++) {
                    lo[k] = lo[k] < p[k] ? lo[k] : p[k];
                    hi[k] = hi[k] > p[k] ? hi[k] : p[k];
                }
                vcube* c = &cubes[id[key]];
                if (c->n + 4 > c->cap) {
                    c->cap = c->cap ? 2 * c->cap : 256;
                    c->v = (uint8_t*)realloc(c->v, c->cap);
                }
                c->v[c->n++] = (uint8_t)(vX % M);
                c->v[c->n++] = (uint8_t)(vY % M);
                c->v[c->n++] = (uint8_t)(vZ % M);
                c->v[c->n++] = 1;
            }
    sink s = {out, cap, 0};
    put32(&s, 0x20584F56); /* "VOX " */
    put32(&s, 150);
    put32(&s, 0x4E49414D); /* MAIN */
    put32(&s, 0);
    put32(&s, 0); /* child size, patched below */
    for (size_t i = 0; i < ncubes; i++) {
        put32(&s, 0x455A4953); /* SIZE */
        put32(&s, 12);
        put32(&s, 0);
        put32(&s, (int32_t)M);
        put32(&s, (int32_t)M);
        put32(&s, (int32_t)M);
        put32(&s, 0x495A5958); /* XYZI */
        put32(&s, (int32_t)(4 * (1 + cubes[i].n / 4)));
        put32(&s, 0);
        put32(&s, (int32_t)(cubes[i].n / 4));
        put(&s, cubes[i].v, cubes[i].n);
    }
    /* root transform: node 0 -> group node 1 */
    put32(&s, 0x4E52546E); put32(&s, 28); put32(&s, 0);
    put32(&s, 0); put32(&s, 0); put32(&s, 1); put32(&s, -1); put32(&s, -1); put32(&s, 1); put32(&s, 0);
    /* root group */
    put32(&s, 0x5052476E); put32(&s, (int32_t)(4 * (2 + ncubes) + 4)); put32(&s, 0);
    put32(&s, 1); put32(&s, 0); put32(&s, (int32_t)ncubes);
    for (size_t i = 0; i < ncubes; i++) put32(&s, (int32_t)(2 + 2 * i));
    const double size[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    for (size_t i = 0; i < ncubes; i++) {
        const vcube* c = &cubes[i];
        const int32_t tx = x86_int(floor((double)(((float)(size_t)(c->ox - minCubeX) + 0.5f) * (float)M) - lo[0] -
                                         size[0] * 0.5));
        const int32_t ty = x86_int(floor((double)(((float)(size_t)(c->oy - minCubeY) + 0.5f) * (float)M) - lo[1] -
                                         size[1] * 0.5));
        const int32_t tz = x86_int((double)floorf(((float)(size_t)(c->oz - minCubeZ) + 0.5f) * (float)M));
        char t[64];
        snprintf(t, sizeof t, "%d %d %d", tx, ty, tz);
        const int32_t tl = (int32_t)strlen(t);
        /* nTRN: 5 ints + empty attribs (4) + frame dict (4 + 4+2 + 4+tl) */
        put32(&s, 0x4E52546E); put32(&s, 20 + 4 + 4 + 6 + 4 + tl); put32(&s, 0);
        put32(&s, (int32_t)(2 + 2 * i)); put32(&s, 0); put32(&s, (int32_t)(3 + 2 * i)); put32(&s, -1);
        put32(&s, 0); put32(&s, 1);
        put32(&s, 1); putstr(&s, "_t"); putstr(&s, t);
        /* nSHP: node, attribs, numModels, model id + {"_f": "0"} */
        put32(&s, 0x5048536E); put32(&s, 8 + 4 + 4 + 4 + 6 + 5); put32(&s, 0);
        put32(&s, (int32_t)(3 + 2 * i)); put32(&s, 0); put32(&s, 1);
        put32(&s, (int32_t)i); put32(&s, 1); putstr(&s, "_f"); putstr(&s, "0");
    }
    put32(&s, 0x41424752); put32(&s, 1024); put32(&s, 0);
    put32(&s, 0x00FFFFFF);
    for (int k = 1; k < 256; k++) put32(&s, 0);
    if (s.len <= s.cap) {
        const uint32_t child = (uint32_t)(s.len - 20);
        for (int k = 0; k < 4; k++) out[16 + k] = (uint8_t)(child >> (8 * k));
    }
    for (size_t i = 0; i < ncubes; i++) free(cubes[i].v);
    free(cubes);
    free(id);
    return s.len;
}
