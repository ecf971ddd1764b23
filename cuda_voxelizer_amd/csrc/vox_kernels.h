// vox_kernels.h -- internal launch functions shared by the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cuda_voxelizer_amd/vox_capi.h"

namespace voxamd {

hipError_t launch_surface_direct(const vox_info& info, const float* tris, uint32_t* table, int layout,
                                 hipStream_t stream);
hipError_t launch_solid_direct(const vox_info& info, const float* tris, uint32_t* table, int layout,
                               unsigned long long* offrow, hipStream_t stream);

}  // namespace voxamd
