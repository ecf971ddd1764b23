// vox_host.h -- host-side helpers shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/cuda_voxelizer_amd/vox_capi.h"

namespace voxamd {

constexpr uint32_t kMaxGrid = 8192;                 // 8192^3 bits = 64 GiB table
constexpr uint64_t kMaxTriangles = 0x7FFFFFFFull;   // triangle ids are int32 in the bins

vox_status fail(vox_status s, const std::string& msg);
vox_status hip_fail(hipError_t e, const char* what);
vox_status validate(const vox_info* info, const float* tris, const uint32_t* table, int layout);

}  // namespace voxamd
