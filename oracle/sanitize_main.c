/*
 * sanitize_main.c -- driver for the sanitizer builds of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY (like the rest of oracle/): built by sanitize.mk
 * into oracle/_san/ under AddressSanitizer + UndefinedBehaviorSanitizer and
 * under ThreadSanitizer, and run by tests/test_oracle_sanitizers.py on the CPU.
 *
 * It drives every exported oracle routine over inputs that reach the
 * boundary code: a closed jittered torus, random and degenerate triangles
 * (points, segments, triangles reaching past the grid so their voxel boxes
 * are clamped), and solid triangles in planes x = c beyond the grid (runs that
 * continue into the next rows and are clipped at the table end), in front of
 * it (empty runs) and over the last rows.  For every grid, layout and mode it
 * checks that the reference-structured mode (one critical section per bit) on
 * one thread, the same mode on four threads, and the lock-free mode on four
 * threads give the same table and the same out-of-row count -- OR and XOR
 * commute, so any difference is a race -- and it sizes every writer's output
 * with the size query and writes into a buffer of exactly that size.
 * Exit status 0 when every check holds.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    float bmin[3];
    float bmax[3];
    uint32_t grid[3];
    uint32_t _pad0;
    uint64_t n_triangles;
    float unit[3];
    uint32_t _pad1;
} voxo_info;

double voxo_surface(const voxo_info* info, float* verts, int64_t nverts, const int32_t* faces, uint32_t* table,
                    int morton, int mode);
double voxo_solid(const voxo_info* info, float* verts, int64_t nverts, const int32_t* faces, uint32_t* table,
                  int morton, int mode, int64_t* out_offrow);
void voxo_make_info(const float* bmin3, const float* bmax3, uint32_t grid, uint64_t n_triangles, voxo_info* out);
void voxo_mesh_bbox(const float* verts, int64_t nverts, float* out_min3, float* out_max3);
void voxo_bbox_cube(const float* in_min3, const float* in_max3, float* out_min3, float* out_max3);
void voxo_set_threads(int n);
uint64_t voxo_binvox_body(const uint32_t* table, uint32_t grid, int morton, uint8_t* out, uint64_t cap);
uint64_t voxo_obj_points(const uint32_t* table, uint32_t grid, int morton, uint8_t* out, uint64_t cap);
uint64_t voxo_obj_cubes(const uint32_t* table, uint32_t grid, int morton, uint8_t* out, uint64_t cap);
uint64_t voxo_vox_file(const uint32_t* table, uint32_t grid, int morton, uint8_t* out, uint64_t cap);

enum { MODE_REF = 0, MODE_FAST = 1 };

typedef struct {
    float* v;
    int32_t* f;
    int64_t nv, nf, cap_v, cap_f;
} mesh;

static uint32_t lcg_state = 12345u;
static float urand(void) {
    lcg_state = lcg_state * 1664525u + 1013904223u;
    return (float)(lcg_state >> 8) * (1.0f / 16777216.0f);
}

static int64_t add_vertex(mesh* m, float x, float y, float z) {
    if (m->nv == m->cap_v) {
        m->cap_v = m->cap_v ? 2 * m->cap_v : 1024;
        m->v = (float*)realloc(m->v, (size_t)m->cap_v * 3 * sizeof(float));
    }
    m->v[3 * m->nv + 0] = x;
    m->v[3 * m->nv + 1] = y;
    m->v[3 * m->nv + 2] = z;
    return m->nv++;
}

static void add_face(mesh* m, int64_t a, int64_t b, int64_t c) {
    if (m->nf == m->cap_f) {
        m->cap_f = m->cap_f ? 2 * m->cap_f : 1024;
        m->f = (int32_t*)realloc(m->f, (size_t)m->cap_f * 3 * sizeof(int32_t));
    }
    m->f[3 * m->nf + 0] = (int32_t)a;
    m->f[3 * m->nf + 1] = (int32_t)b;
    m->f[3 * m->nf + 2] = (int32_t)c;
    m->nf++;
}

static void add_tri(mesh* m, const float* p) {
    int64_t a = add_vertex(m, p[0], p[1], p[2]);
    int64_t b = add_vertex(m, p[3], p[4], p[5]);
    int64_t c = add_vertex(m, p[6], p[7], p[8]);
    add_face(m, a, b, c);
}

/* closed torus around the grid centre, radii jittered by up to 2% */
static void torus(mesh* m, float G, int nu, int nw) {
    const float R = 0.3f * G, r = 0.12f * G, c = 0.5f * G;
    int64_t base = m->nv;
    for (int i = 0; i < nu; i++) {
        for (int j = 0; j < nw; j++) {
            float u = 6.2831853f * (float)i / (float)nu, w = 6.2831853f * (float)j / (float)nw;
            float rr = r * (1.0f + 0.02f * (urand() - 0.5f));
            float x = (R + rr * cosf(w)) * cosf(u), y = (R + rr * cosf(w)) * sinf(u), z = rr * sinf(w);
            add_vertex(m, c + x, c + 0.9f * y, c + z);
        }
    }
    for (int i = 0; i < nu; i++) {
        for (int j = 0; j < nw; j++) {
            int64_t a = base + i * nw + j, b = base + ((i + 1) % nu) * nw + j;
            int64_t d = base + i * nw + (j + 1) % nw, e = base + ((i + 1) % nu) * nw + (j + 1) % nw;
            add_face(m, a, b, e);
            add_face(m, a, e, d);
        }
    }
}

/* surface extras: random triangles, degenerate ones, ones reaching past the grid */
static void surface_extras(mesh* m, float G) {
    for (int k = 0; k < 200; k++) {
        float p[9];
        float cx = urand() * G, cy = urand() * G, cz = urand() * G, s = 0.5f + 4.0f * urand();
        for (int q = 0; q < 3; q++) {
            p[3 * q + 0] = cx + s * (urand() - 0.5f);
            p[3 * q + 1] = cy + s * (urand() - 0.5f);
            p[3 * q + 2] = cz + s * (urand() - 0.5f);
        }
        add_tri(m, p);
    }
    const float pt[9] = {3.5f, 4.5f, 5.5f, 3.5f, 4.5f, 5.5f, 3.5f, 4.5f, 5.5f};
    const float seg[9] = {1.2f, 1.3f, 1.4f, 9.7f, 5.1f, 3.3f, 1.2f, 1.3f, 1.4f};
    const float past[9] = {G - 2.5f, G - 1.5f, 0.5f, G + 3.0f, G - 0.5f, 2.5f, G - 1.0f, G + 2.0f, -1.5f};
    const float under[9] = {-2.0f, 0.5f, 0.5f, 1.5f, -1.0f, 0.7f, 0.5f, 1.5f, -0.5f};
    add_tri(m, pt);
    add_tri(m, seg);
    add_tri(m, past);
    add_tri(m, under);
}

/* solid extras: x-plane triangles with runs beyond, inside and before the grid */
static void solid_extras(mesh* m, float G) {
    const float cs[4] = {G + 1.7f, 2.0f * G + 0.2f, 0.2f, -0.7f};
    const float yz[3][6] = {{0.2f, 0.3f, 3.1f, 0.6f, 0.9f, 3.4f},
                            {0.2f, 0.3f, G - 0.9f, 0.6f, 0.9f, G - 0.6f},
                            {G - 2.8f, G - 2.7f, G - 0.1f, G - 2.4f, G - 2.1f, G - 0.1f}};
    for (int i = 0; i < 4; i++) {
        for (int j = 0; j < 3; j++) {
            float p[9] = {cs[i], yz[j][0], yz[j][1], cs[i], yz[j][2], yz[j][3], cs[i], yz[j][4], yz[j][5]};
            add_tri(m, p);
        }
    }
}

static int failures = 0, checks = 0;

static void expect(int ok, const char* what, uint32_t G, int solid, int morton) {
    checks++;
    if (!ok) {
        failures++;
        fprintf(stderr, "FAILED: %s (G=%u solid=%d morton=%d)\n", what, G, solid, morton);
    }
}

static uint32_t* run(const voxo_info* info, const mesh* m, size_t words, int solid, int morton, int mode,
                     int threads, int64_t* offrow) {
    uint32_t* t = (uint32_t*)calloc(words, sizeof(uint32_t));
    float* v = (float*)malloc((size_t)m->nv * 3 * sizeof(float));  /* the prepass translates in place */
    memcpy(v, m->v, (size_t)m->nv * 3 * sizeof(float));
    voxo_set_threads(threads);
    *offrow = 0;
    if (solid)
        voxo_solid(info, v, m->nv, m->f, t, morton, mode, offrow);
    else
        voxo_surface(info, v, m->nv, m->f, t, morton, mode);
    free(v);
    return t;
}

typedef uint64_t (*writer_fn)(const uint32_t*, uint32_t, int, uint8_t*, uint64_t);

static void writers(const uint32_t* t, uint32_t G, int solid, int morton) {
    writer_fn fns[4] = {voxo_binvox_body, voxo_obj_points, voxo_obj_cubes, voxo_vox_file};
    const char* names[4] = {"binvox body", "obj points", "obj cubes", "vox file"};
    for (int k = 0; k < 4; k++) {
        uint64_t n = fns[k](t, G, morton, NULL, 0);
        uint8_t* out = (uint8_t*)malloc(n ? n : 1);
        uint64_t n2 = fns[k](t, G, morton, out, n);
        expect(n2 == n && n > 0, names[k], G, solid, morton);
        free(out);
    }
}

int main(void) {
    const uint32_t grids[3] = {32, 48, 64};
    for (int gi = 0; gi < 3; gi++) {
        const uint32_t G = grids[gi];
        for (int solid = 0; solid < 2; solid++) {
            mesh m = {0};
            torus(&m, (float)G, 40, 20);
            if (solid)
                solid_extras(&m, (float)G);
            else
                surface_extras(&m, (float)G);
            /* the grid box [0, G]^3 (unit 1): the extras reach past it on purpose */
            const float bmin[3] = {0.0f, 0.0f, 0.0f}, bmax[3] = {(float)G, (float)G, (float)G};
            voxo_info info;
            voxo_make_info(bmin, bmax, G, (uint64_t)m.nf, &info);
            const size_t words = ((size_t)G * G * G + 31) / 32;
            for (int morton = 0; morton < 2; morton++) {
                if (morton && (G & (G - 1))) continue;  /* Morton tables need a power-of-two grid */
                int64_t o1, o4, of;
                uint32_t* ref1 = run(&info, &m, words, solid, morton, MODE_REF, 1, &o1);
                uint32_t* ref4 = run(&info, &m, words, solid, morton, MODE_REF, 4, &o4);
                uint32_t* fast4 = run(&info, &m, words, solid, morton, MODE_FAST, 4, &of);
                expect(memcmp(ref1, ref4, words * 4) == 0, "reference mode, 1 vs 4 threads", G, solid, morton);
                expect(memcmp(ref1, fast4, words * 4) == 0, "reference vs lock-free mode", G, solid, morton);
                expect(o1 == o4 && o1 == of, "out-of-row counts", G, solid, morton);
                size_t set = 0;
                for (size_t w = 0; w < words; w++) set += (size_t)__builtin_popcount(ref1[w]);
                expect(set > 0, "non-empty table", G, solid, morton);
                if (solid) expect(o1 > 0, "out-of-row runs reached", G, solid, morton);
                writers(fast4, G, solid, morton);
                free(ref1);
                free(ref4);
                free(fast4);
            }
            /* host prep over the same vertices */
            float mn[3], mx[3], cmn[3], cmx[3];
            voxo_mesh_bbox(m.v, m.nv, mn, mx);
            voxo_bbox_cube(mn, mx, cmn, cmx);
            expect(cmx[0] - cmn[0] >= mx[0] - mn[0], "bbox cube", G, solid, 0);
            free(m.v);
            free(m.f);
        }
    }
    printf("oracle sanitizer run: %d checks, %d failed\n", checks, failures);
    return failures ? 1 : 0;
}
