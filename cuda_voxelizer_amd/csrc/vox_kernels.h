// vox_kernels.h -- internal launch functions shared by the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cuda_voxelizer_amd/vox_capi.h"
#include "tile.h"

namespace voxamd {

hipError_t launch_surface_direct(const vox_info& info, const float* tris, uint32_t* table, int layout,
                                 hipStream_t stream);
hipError_t launch_solid_direct(const vox_info& info, const float* tris, uint32_t* table, int layout,
                               unsigned long long* offrow, hipStream_t stream);

bool make_plan(int G, int layout, bool solid, TilePlan& P);
WorkspaceLayout workspace_layout(uint64_t tiles_cap, uint64_t n_cap, int max_tiles_per_tri);
hipError_t launch_tiled(const vox_info& info, const float* tris, uint32_t* table, const TilePlan& P, bool solid,
                        int layout, bool accumulate, void* ws, const WorkspaceLayout& L,
                        unsigned long long* offrow, hipStream_t stream);

bool chunk_supported(uint32_t G);
hipError_t launch_fill_zero(uint32_t* table, uint64_t words, hipStream_t stream);
hipError_t launch_surface_chunk(const vox_info& info, const float* tris, uint32_t* table, int layout,
                                hipStream_t stream);

}  // namespace voxamd
