// dropin_caller.cpp -- the reference's own call sequence, written against
// <cuda_voxelizer_amd/voxelize.h> the way main.cpp uses voxelize.cu:
//   bbox over the mesh (trimesh2 need_bbox, main.cpp:179)
//   createMeshBBCube<float3>(AABox<float3>(min, max))     main.cpp:184
//   voxinfo(cube, make_uint3(G, G, G), n_faces)             main.cpp:186
//   vtable_size = ceil(gx*gy*gz / 32.0f) * 4               main.cpp:190
//   triangles packed v0 v1 v2 per face (meshToGPU_managed)  main.cpp:61-80
//   voxelize(...) / voxelize_solid(...)                     main.cpp:218, 221
// Device buffers are hipMalloc'd (the reference uses managed memory, which
// the drop-in accepts too) and the table is zeroed before the call, as the
// reference's fresh managed pages are.  Test code: built into
// libdropin_caller.so and driven from Python through dropin_run().
#include <cuda_voxelizer_amd/voxelize.h>

#include <cmath>
#include <cstring>
#include <vector>

extern "C" int dropin_run(const float* vertices, int n_vertices, const int* faces, int n_faces, int gridsize,
                          int solid, int morton, int five_arg_form, unsigned int* table_out, float* info_out) {
    if (n_vertices <= 0 || n_faces < 0 || gridsize <= 0) return 1;
    // trimesh2 need_bbox: component-wise min and max over all vertices
    float3 mn = make_float3(vertices[0], vertices[1], vertices[2]), mx = mn;
    for (int i = 1; i < n_vertices; i++) {
        const float x = vertices[3 * i], y = vertices[3 * i + 1], z = vertices[3 * i + 2];
        mn.x = x < mn.x ? x : mn.x; mn.y = y < mn.y ? y : mn.y; mn.z = z < mn.z ? z : mn.z;
        mx.x = x > mx.x ? x : mx.x; mx.y = y > mx.y ? y : mx.y; mx.z = z > mx.z ? z : mx.z;
    }
    AABox<float3> bbox_mesh_cubed = createMeshBBCube<float3>(AABox<float3>(mn, mx));
    voxinfo voxelization_info(bbox_mesh_cubed, make_uint3(gridsize, gridsize, gridsize), (size_t)n_faces);
    const size_t vtable_size = static_cast<size_t>(
        ceil(static_cast<size_t>(voxelization_info.gridsize.x) * static_cast<size_t>(voxelization_info.gridsize.y) *
             static_cast<size_t>(voxelization_info.gridsize.z) / 32.0f) * 4);
    // meshToGPU_managed: v0.xyz v1.xyz v2.xyz per face
    std::vector<float> tris((size_t)n_faces * 9);
    for (int f = 0; f < n_faces; f++)
        for (int k = 0; k < 3; k++)
            for (int c = 0; c < 3; c++) tris[(size_t)f * 9 + 3 * k + c] = vertices[3 * faces[3 * f + k] + c];
    float* device_triangles = nullptr;
    unsigned int* vtable = nullptr;
    if (hipMalloc((void**)&device_triangles, tris.size() * sizeof(float) + 4) != hipSuccess) return 2;
    if (hipMalloc((void**)&vtable, vtable_size) != hipSuccess) return 2;
    if (hipMemcpy(device_triangles, tris.data(), tris.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        return 3;
    if (hipMemset(vtable, 0, vtable_size) != hipSuccess) return 3;
    voxelize_set_verbose(false);
    if (solid)
        voxelize_solid(voxelization_info, device_triangles, vtable, morton != 0);
    else if (five_arg_form)
        voxelize(voxelization_info, device_triangles, vtable, /*useThrustPath=*/false, morton != 0);
    else
        voxelize(voxelization_info, device_triangles, vtable, morton != 0);
    if (hipMemcpy(table_out, vtable, vtable_size, hipMemcpyDeviceToHost) != hipSuccess) return 4;
    std::memcpy(info_out, &voxelization_info, sizeof(voxelization_info));  // the 64-byte voxinfo
    (void)hipFree(device_triangles);
    (void)hipFree(vtable);
    return 0;
}
