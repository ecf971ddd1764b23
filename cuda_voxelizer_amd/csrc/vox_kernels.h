// vox_kernels.h -- internal launch functions shared by the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cuda_voxelizer_amd/vox_capi.h"
#include "region.h"
#include "slab.h"
#include "tile.h"

namespace voxamd {

hipError_t launch_surface_direct(const vox_info& info, const float* tris, uint32_t* table, int layout,
                                 hipStream_t stream);
hipError_t launch_solid_direct(const vox_info& info, const float* tris, uint32_t* table, int layout,
                               unsigned long long* offrow, hipStream_t stream);

uint32_t direct_units(uint64_t n_triangles);
hipError_t launch_direct_jobs(const ItemDev& one, const ItemMap& M, uint32_t units, bool solid, hipStream_t stream);
bool make_plan(int G, int layout, bool solid, TilePlan& P, int scale = 0);
WorkspaceLayout workspace_layout(uint64_t tiles_cap, uint64_t n_cap, int max_tiles_per_tri);
hipError_t launch_tiled_jobs(const ItemDev& one, const ItemMap& M_bin, const ItemMap& M_tile, const ItemMap& M_tri,
                             uint32_t total_bin_blocks, uint32_t total_tiles, bool solid, uint32_t max_tile_words,
                             void* ws, const WorkspaceLayout& L, hipStream_t stream);
hipError_t launch_tiled(const vox_info& info, const float* tris, uint32_t* table, const TilePlan& P, bool solid,
                        int layout, bool accumulate, void* ws, const WorkspaceLayout& L,
                        unsigned long long* offrow, hipStream_t stream);

bool marker_supported(uint32_t G, int layout, bool accumulate);
bool marker_batch_supported(uint32_t G, int layout, bool accumulate);
bool layout_convert_fast(uint32_t G);
hipError_t launch_layout_convert(const uint32_t* src, uint32_t* dst, uint32_t G, bool to_morton, hipStream_t stream);
uint64_t marker_ws_bytes();
uint32_t marker_block_units(uint64_t n_triangles);
uint32_t marker_row_units(uint32_t G);
hipError_t launch_marker_jobs(const ItemMap& m_blocks, const ItemMap& m_rows, const ItemMap& m_tris,
                              uint32_t blocks, uint32_t row_units, void* mws, hipStream_t stream);
hipError_t launch_solid_marker(const vox_info& info, const float* tris, uint32_t* table, int layout,
                               bool accumulate, unsigned long long* offrow, hipStream_t stream, uint32_t big = 0);
// The allocation-free solid form: its scratch is a static device array of
// static_scratch_bytes() bytes per GPU (address taken once per device).
size_t static_scratch_bytes();
hipError_t static_scratch_init(int device, void** out);
hipError_t launch_solid_marker_static(const vox_info& info, const float* tris, uint32_t* table,
                                      unsigned long long* offrow, hipStream_t stream);

bool make_slab_plan(uint32_t G, SlabPlan& P, int lg_ty = -1, int lg_tz = -1, int split_lg = -1, int wg = 0);
SlabLayout slab_layout(uint64_t tiles_cap, uint64_t n_cap);
bool slab_layout_fits(const SlabPlan& P, uint64_t n, const SlabLayout& L);
hipError_t launch_slab(const vox_info& info, const float* tris, uint32_t* table, const SlabPlan& P, bool solid,
                       bool accumulate, unsigned long long* offrow, void* mem, const SlabLayout& L,
                       hipStream_t stream);

// Region algorithm (region.hip): phase 1 stages (word, mask) pairs per output
// region (surface: chunk.hip, solid: marker.hip), the build pass writes each
// region once.  K: capacity overrides (flag bits), -1 fields are built-in.
bool make_region_plan(uint32_t G, int layout, bool solid, RegionPlan& P, const RegionKnobs& K);
bool region_supported(uint32_t G, int layout, bool solid);
// workgroups: phase-1 workgroups of the call (each stages into its own slice)
hipError_t region_workspace(const RegionPlan& P, uint64_t workgroups, hipStream_t stream, RegionWs& W);
hipError_t launch_region_build(const RegionPlan& P, const RegionWs& W, uint32_t* table, bool accumulate,
                               hipStream_t stream);
hipError_t launch_surface_region_chunks(const vox_info& info, const float* tris, int layout, const RegionArgs& RA,
                                        hipStream_t stream, uint32_t big, int waves);
hipError_t launch_surface_region(const vox_info& info, const float* tris, uint32_t* table, int layout,
                                 bool accumulate, hipStream_t stream, uint32_t big = 0, int waves = 0,
                                 const RegionKnobs& K = RegionKnobs{});
hipError_t launch_solid_region(const vox_info& info, const float* tris, uint32_t* table, int layout,
                               bool accumulate, unsigned long long* offrow, hipStream_t stream, uint32_t big = 0,
                               const RegionKnobs& K = RegionKnobs{});
// VOX_REGION_* bits of the flags
inline RegionKnobs region_knobs(unsigned flags) {
    RegionKnobs K;
    const unsigned w = (flags >> 8) & 15u, st = (flags >> 12) & 31u, d = (flags >> 17) & 15u;
    if (w) K.lg_words = (int)w + 9;
    if (st) K.lg_stage = (int)st - 1;
    if (d) K.lg_desc = (int)d - 1;
    return K;
}

bool chunk_supported(uint32_t G);
// A coarse mesh has few triangles for its grid, so each covers a large voxel
// box and many columns.  Solid default: the output-tiled algorithm for a
// coarse mesh, which spreads its columns over the whole device (1,280
// triangles at 256^3: tile 0.09 ms vs marker 0.36 ms), the marker algorithm
// otherwise (1,310,720 at 1024^3: marker 0.14 ms vs tile 0.41 ms).  Surface
// default: chunk for every mesh; a coarse mesh's chunk splits its queued
// items over up to 32 workgroups (chunk_parts), which measured faster than
// tile from 6 to 1,310,720 triangles at 64^3 to 2048^3 (1,280 at 256^3: 0.057
// vs 0.112 ms; a 12-triangle cube at 2048^3: 2.8 vs 8.4 ms), except for a few
// hundred triangles at 128^3 to 512^3 (at most 1.5x slower).
inline bool coarse_mesh(uint64_t n_triangles, uint32_t G) {
    return n_triangles < (uint64_t)G * G / 4u;
}
// Solid default on small grids: a mesh of at least 4 G^2 triangles at G <= 64
// has boxes of a few columns each, and one lane per triangle with its
// atomics straight to the table beats the marker pass and its suffix (65,536
// triangles at 64^3: direct 0.0098 ms vs marker 0.0161 ms; 5,120 at 32^3:
// 0.0106 vs 0.0165 ms); at 128^3, or below 4 G^2, marker is faster (81,920 at
// 128^3: 0.0133 vs 0.0169 ms; the bunny's 5,110 at 64^3: 0.032 vs 0.046 ms).
inline bool prefer_direct_solid(uint64_t n_triangles, uint32_t G) {
    return G <= 64u && n_triangles >= 4ull * G * G;
}
hipError_t launch_fill_zero(const vox_info& info, uint32_t* table, hipStream_t stream);
uint32_t fill_units(uint32_t G);
uint32_t chunk_span(uint32_t G, uint64_t n_triangles);
uint32_t chunk_parts(uint32_t G, uint64_t n_triangles);
uint32_t chunk_units(uint64_t n_triangles, uint32_t span, uint32_t parts);
hipError_t launch_fill_jobs(const ItemDev& one, const ItemMap& M, uint32_t units, hipStream_t stream);
// What a chunk kernel instantiation knows of its jobs' layout: nothing, linear
// with G a multiple of 32 (whole words per row), or Morton with G <= 1024
// (codes fit 30 bits).
enum ChunkSpec : int { kSpecAny = 0, kSpecLinear = 1, kSpecMorton = 2 };
int chunk_spec(int layout, uint32_t G);
hipError_t launch_chunk_jobs(const ItemDev& one, const ItemMap& M, uint32_t units, bool loop, hipStream_t stream,
                             int waves = 0, int spec = kSpecAny);
hipError_t launch_surface_chunk(const vox_info& info, const float* tris, uint32_t* table, int layout,
                                hipStream_t stream, uint32_t big = 0, int waves = 0);
// Tuning bits of the flags (VOX_TUNE_*): queue threshold, chunk waves per SIMD, tile size.
inline uint32_t tune_big(unsigned flags) {
    const unsigned lg = (flags >> 24) & 15u;
    return lg ? (1u << (lg - 1u)) : 0u;
}
inline int tune_waves(unsigned flags) {
    const unsigned w = (flags >> 28) & 3u;
    return w == 1u ? 5 : (w == 2u ? 6 : 0);
}
inline int tune_tile_scale(unsigned flags) {
    const unsigned t = (flags >> 30) & 3u;
    return t == 1u ? -1 : (t == 3u ? 1 : 0);
}

}  // namespace voxamd
