// voxelize.h -- C++ drop-in for the reference's host API (Forceflow/cuda_voxelizer).
//
// A program written against the reference includes this header instead of
// forward-declaring the CUDA entry points (main.cpp:22-24) and links
// libvoxelizer_amd.so.  Types and semantics follow the reference:
//   AABox<T>            util.h:41-47
//   voxinfo             util.h:50-69 (64-byte layout, unit = (max-min)/float(G))
//   createMeshBBCube    util.h:80-110
//   voxelize(...)       voxelize.cu:192-238   (surface, OR into vtable)
//   voxelize_solid(...) voxelize_solid.cu:147-193 (solid, XOR into vtable)
// Both entry points are synchronous, run on the default stream, OR/XOR into
// the caller's (pre-zeroed) table exactly like the reference, and on a HIP
// error print "[Err] ..." and exit(EXIT_FAILURE) like checkCudaErrors
// (helper_cuda.h:566-579).  `useThrustPath` is accepted for source
// compatibility with newer reference releases and has no effect on the result.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdio>

#include "vox_capi.h"

template <typename T>
struct AABox {
    T min;
    T max;
    __host__ __device__ AABox() : min(T()), max(T()) {}
    __host__ __device__ AABox(T mn, T mx) : min(mn), max(mx) {}
};

struct voxinfo {
    AABox<float3> bbox;
    uint3 gridsize;
    size_t n_triangles;
    float3 unit;

    voxinfo(const AABox<float3> b, const uint3 g, const size_t n)
        : bbox(b), gridsize(g), n_triangles(n) {
        unit.x = (bbox.max.x - bbox.min.x) / float(gridsize.x);
        unit.y = (bbox.max.y - bbox.min.y) / float(gridsize.y);
        unit.z = (bbox.max.z - bbox.min.z) / float(gridsize.z);
    }

    void print() const {
        std::fprintf(stdout, "[Voxelization] Bounding Box: (%f,%f,%f)-(%f,%f,%f) \n", bbox.min.x,
                     bbox.min.y, bbox.min.z, bbox.max.x, bbox.max.y, bbox.max.z);
        std::fprintf(stdout, "[Voxelization] Grid size: %i %i %i \n", gridsize.x, gridsize.y,
                     gridsize.z);
        std::fprintf(stdout, "[Voxelization] Triangles: %zu \n", n_triangles);
        std::fprintf(stdout, "[Voxelization] Unit length: x: %f y: %f z: %f\n", unit.x, unit.y,
                     unit.z);
    }
};

static_assert(sizeof(voxinfo) == sizeof(vox_info), "voxinfo must keep the reference's 64-byte layout");
static_assert(offsetof(voxinfo, gridsize) == offsetof(vox_info, gridsize), "gridsize offset");
static_assert(offsetof(voxinfo, n_triangles) == offsetof(vox_info, n_triangles), "n_triangles offset");
static_assert(offsetof(voxinfo, unit) == offsetof(vox_info, unit), "unit offset");

// createMeshBBCube for float3 boxes (util.h:80-110), evaluated by the library.
AABox<float3> createMeshBBCube(AABox<float3> box);

// The reference's template form, called as createMeshBBCube<float3>(box)
// (util.h:80, main.cpp:184); only float3 boxes are voxelized.
template <typename T>
AABox<T> createMeshBBCube(AABox<T> box);
template <>
inline AABox<float3> createMeshBBCube<float3>(AABox<float3> box) {
    return createMeshBBCube(box);  // the non-template overload above
}

// Surface voxelization: reference v0.6 form and the 5-argument form.
void voxelize(const voxinfo& v, float* triangle_data, unsigned int* vtable, bool morton_code);
void voxelize(const voxinfo& v, float* triangle_data, unsigned int* vtable, bool useThrustPath,
              bool morton_code);

// Solid (parity) voxelization.
void voxelize_solid(const voxinfo& v, float* triangle_data, unsigned int* vtable, bool morton_code);

// When true, voxelize/voxelize_solid print "[Perf] Voxelization GPU time: %.1f ms"
// as the reference does (voxelize.cu:222).  Default: true.
void voxelize_set_verbose(bool verbose);
