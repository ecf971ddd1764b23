// tile.h -- output tiling shared by the binned (tile-owned) kernels.
//
// The bit table is split into tiles that are each a set of whole words; one
// workgroup owns a tile, builds it in LDS and writes every one of its words
// exactly once.  Tile shapes:
//   linear layout (G % 32 == 0): all x (G bits per row) x ty rows in y x tz
//     slabs in z; a tile is tz contiguous segments of ty*G/32 words.
//   Morton layout, surface (G power of two): an aligned cube of edge T, whose
//     T^3/32 words are one contiguous range; tile id = Morton code of the cube.
//   Morton layout, solid: all x x T (y) x T (z); the parity fill needs whole
//     x rows.  Output = G/T aligned cubes of T^3/32 contiguous words.
#pragma once

#include <stdint.h>

#include "../../include/cuda_voxelizer_amd/vox_capi.h"

namespace voxamd {

enum TileKind : int { TILE_LINEAR = 0, TILE_MORTON_CUBE = 1, TILE_MORTON_ROWS = 2 };

struct TilePlan {
    int kind;
    int G;            // grid edge
    int tx, ty, tz;   // tile extent in voxels (tx == G for row tiles)
    int ntx, nty, ntz;
    int ntiles;
    int tile_words;   // LDS words per tile
    int row_words;    // G / 32 for row tiles
    int max_tiles_per_tri;  // triangles spanning more tiles go to the wide path
    int lg_cube;      // log2(tx) for Morton cubes
    int pad;
};

// One voxelization job as the kernels see it.
struct ItemDev {
    vox_info info;
    const float* tris;
    uint32_t* table;
    unsigned long long* offrow;  // solid: out-of-row run counter (may be null)
    TilePlan P;
    int32_t layout, solid, accumulate;
    uint32_t tile_base;  // first global tile id of this job (tile algorithm)
    uint64_t tri_base;   // first global triangle id of this job
    uint32_t span;       // chunk algorithm: chunks per workgroup (0 or 1: one)
    uint32_t parts;      // chunk algorithm: workgroups per chunk (0 or 1: one)
    uint32_t big;        // chunk: voxel boxes above this are shared by the workgroup; marker: columns
                         // above this (0: the built-in threshold)
    uint32_t pad2;
};

// Work mapping of one launch: a single job passed by value (items == nullptr)
// or a device array of jobs with base[i] = first work unit (block, tile, ...)
// of job i in this launch and base[nitems] = total units.
struct ItemMap {
    const ItemDev* items;
    const uint32_t* base;
    int nitems;
    int pad;
};

// bytes of device workspace a call with this plan needs
struct WorkspaceLayout {
    uint64_t counts_off, offsets_off, list_off, wide_off, misc_off, spill_off, total;
    uint64_t list_cap, wide_cap, spill_cap;
};

}  // namespace voxamd
