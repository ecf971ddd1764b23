// Host fuzz driver for vox_load_mesh (tests/test_loader_sanitizers.py builds it
// with g++ under AddressSanitizer + UndefinedBehaviorSanitizer, CPU only).
//
// Each iteration takes a small valid OBJ, OFF, ASCII STL, ASCII PLY or binary
// PLY file, applies 1-6 seeded mutations (byte deletions, insertions of
// boundary tokens -- huge or negative counts and indices, nan, inf, extra PLY
// element and property lines -- random bytes, truncations, random 8-byte
// doubles), writes it and loads it.  A load either fails with a status or
// returns faces whose indices are all in range; the sanitizers stop the run
// on any invalid access, leak or undefined conversion.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>

#include "../../include/cuda_voxelizer_amd/vox_capi.h"
#include "../../cuda_voxelizer_amd/csrc/vox_host.h"

// The loader reports errors through the library's error helpers, which live
// with the device entry points; the driver links the loader alone.
namespace voxamd {
vox_status fail(vox_status s, const std::string&) { return s; }
vox_status hip_fail(hipError_t, const char*) { return VOX_ERR_HIP; }
}  // namespace voxamd

static const char* kSeeds[] = {
    "v 0 0 0\nv 1 0 0\nv 0 1 0\nv 0 0 1\nf 1 2 3\nf 1/1/1 3/2/2 4\nf -1 -2 -3 -4\n",
    "OFF\n4 2 0\n0 0 0\n1 0 0\n0 1 0\n0 0 1\n3 0 1 2\n4 0 1 2 3\n",
    "solid x\nfacet normal 0 0 1\nouter loop\nvertex 0 0 0\nvertex 1 0 0\nvertex 0 1 0\nendloop\nendfacet\n"
    "endsolid\n",
    "ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\nproperty float y\nproperty float z\n"
    "element face 1\nproperty list uchar int vertex_indices\nend_header\n0 0 0\n1 0 0\n0 1 0\n3 0 1 2\n",
    "ply\nformat binary_little_endian 1.0\nelement vertex 3\nproperty double x\nproperty double y\n"
    "property double z\nelement face 1\nproperty list uchar uint vertex_indices\nend_header\n",
};
static const char* kExt[] = {"obj", "off", "stl", "ply", "ply"};
static const char* kTokens[] = {
    " ", "\n", "-", "/", "#", "0", "3", "f", "v", "OFF", "nan", "inf", "1e300", "1e6", "1000001",
    "999999999999", "2147483648", "-2147483649", "4294967295", "9223372036854775807",
    "-9223372036854775808", "6148914691236517206", "end_header\n", "property double x\n",
    "property list uchar int vertex_indices\n", "element face 99999999999\n",
    "element junk 1000000000000000000\n"};

int main(int argc, char** argv) {
    const int iters = argc > 1 ? atoi(argv[1]) : 20000;
    const std::string dir = argc > 2 ? argv[2] : "/tmp";
    std::mt19937_64 rng(0xF0221);
    for (int it = 0; it < iters; it++) {
        const int s = (int)(rng() % 5);
        std::string d = kSeeds[s];
        if (s == 4) {  // binary body: three vertices, one face
            const double v[9] = {0, 0, 0, 1, 0, 0, 0, 1, 0};
            const unsigned char c = 3;
            const unsigned idx[3] = {0, 1, 2};
            d.append((const char*)v, sizeof v);
            d.append((const char*)&c, 1);
            d.append((const char*)idx, sizeof idx);
        }
        const int nmut = 1 + (int)(rng() % 6);
        for (int m = 0; m < nmut; m++) {
            const int op = (int)(rng() % 5);
            const size_t pos = rng() % (d.size() + 1);
            if (op == 0 && pos < d.size()) {
                d.erase(pos, 1 + rng() % 8);
            } else if (op == 1) {
                d.insert(pos, kTokens[rng() % (sizeof kTokens / sizeof *kTokens)]);
            } else if (op == 2 && pos < d.size()) {
                d[pos] = (char)(rng() & 0xFF);
            } else if (op == 3) {
                d.resize(rng() % (d.size() + 1));
            } else if (op == 4) {
                const uint64_t bits = rng();
                d.insert(pos, (const char*)&bits, 8);
            }
        }
        const std::string path = dir + "/fuzz_mesh." + kExt[s];
        FILE* f = fopen(path.c_str(), "wb");
        if (!f) return 3;
        fwrite(d.data(), 1, d.size(), f);
        fclose(f);
        vox_mesh mesh;
        if (vox_load_mesh(path.c_str(), &mesh) == VOX_OK) {
            for (uint64_t i = 0; i < 3 * mesh.n_faces; i++)
                if (mesh.faces[i] < 0 || (uint64_t)mesh.faces[i] >= mesh.n_vertices) {
                    fprintf(stderr, "face index out of range at iteration %d\n", it);
                    return 2;
                }
            vox_free_mesh(&mesh);
        }
    }
    printf("loader fuzz: %d files, 0 failed\n", iters);
    return 0;
}
