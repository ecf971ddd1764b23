// slab.h -- plan and workspace of the write-once slab kernels (slab.hip).
#pragma once

#include <stdint.h>

namespace voxamd {

constexpr int kMaxTiles = 4;                // tiles a binned triangle may touch (else: wide)
constexpr uint32_t kSlabSpillCap = 1u << 16;  // solid runs leaving their row, recorded per call

// Tiles of ty x tz whole rows (all x) of a linear G^3 table.
struct SlabPlan {
    int G;
    int ty, tz;      // rows per tile in y and z
    int nty, ntz;    // tiles per axis
    int ntiles;
    int rw;          // words per row (G / 32)
    int tile_words;  // LDS words per tile
    unsigned split;  // surface: a tile with more entries is cut along x into sub-tiles
    int max_lgp;     // at most 2^max_lgp sub-tiles (rows of at least 4 words)
    int wide_wg;     // raster workgroups of 1024 threads (else 256)
};

// Device workspace of one call (pointers into one allocation).
struct SlabWs {
    unsigned* misc;               // counters (zero between calls)
    unsigned* cnt;                // per-tile counts of hashed pairs (zero between calls)
    unsigned* cntf;               // per-tile counts of pairs no hash slot took (zero between calls)
    unsigned* off;                // per-tile list offsets, [ntiles] = total
    unsigned* cur;                // scatter cursors
    unsigned* key;                // per-triangle tile range
    unsigned* list;               // per-tile triangle ids
    unsigned* units;              // raster work units (tile << 8 | log2(parts) << 4 | part)
    unsigned* brec;               // per binning workgroup: (tile, base, hash slot) of its ranges
    unsigned* brec_n;             // per binning workgroup: number of ranges
    uint64_t brec_stride;         // words of records per binning workgroup
    unsigned* wide;               // triangles touching more than kMaxTiles tiles
    unsigned long long* spill;    // solid: [first bit, last bit] of runs past their row
};

struct SlabLayout {
    uint64_t misc_off, cnt_off, cntf_off, off_off, cur_off, key_off, list_off, units_off, brec_off, brec_n_off;
    uint64_t wide_off, spill_off, total;
    uint64_t brec_stride;
    uint64_t zero_bytes;  // prefix that must be zero before the first call
    uint64_t tiles_cap, n_cap;
};

}  // namespace voxamd
