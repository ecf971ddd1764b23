// slab.hip -- write-once voxelization over slabs of whole x rows (linear layout).
//
// The table is split into tiles of ty x tz whole rows (every x; at 1024^3 a
// tile is 16 x 16 rows = 32 KB), one workgroup per tile.  A call is four
// kernels on one stream and no host synchronisation:
//   1. slab_count   per triangle: its clamped y/z extent -- surface: the voxel
//                   box of cpu_voxelizer.cpp:74-80; solid: the column range of
//                   :280-290 (valid triangles only) -- and so the tiles it
//                   touches (at most kMaxTiles, else the triangle is "wide");
//                   counts merged per workgroup in an LDS hash, then one global
//                   atomicAdd per touched tile; a 4-byte key per triangle for
//                   pass 2.  The last workgroup scans the counts into list
//                   offsets and returns the counts to zero.
//   2. slab_scatter triangle ids into per-tile lists: slots counted in the LDS
//                   hash, one global range reservation per touched tile per
//                   workgroup.
//   3. slab_raster  one workgroup per tile builds the tile in LDS -- surface:
//                   ds_or of every row's passing voxels; solid: one ds_xor
//                   marker per hit column at its run end, then each row
//                   becomes the suffix XOR of its markers -- and writes every
//                   word of the tile once with 16-byte stores.  That write is
//                   also the zero-fill: no fill pass, no global atomics.
//   4. slab_finish  wide triangles, and solid runs that leave their row (the
//                   reference continues them along the linear index into the
//                   next rows, cpu_voxelizer.cpp:296-305), applied to the
//                   finished table with global atomics; the last workgroup
//                   returns the counters to zero, so the workspace needs no
//                   reset between calls.
// Every test value is vox_device.h's, i.e. the reference CPU path's bit for
// bit; only which voxels a lane evaluates, and where the bits are merged,
// differ.
#include "raster.h"
#include "slab.h"
#include "vox_device.h"
#include "vox_kernels.h"

#pragma clang fp contract(off)

namespace voxamd {

constexpr int kSBlock = 256;
constexpr int kHash = 512;            // LDS hash slots (tile -> count) of the binning passes
constexpr int kBinIters = 8;          // rounds of kSBlock triangles per binning workgroup
constexpr int kBinChunk = kSBlock * kBinIters;
constexpr int kDirectTiles = 8192;    // at most this many tiles: per-tile LDS counters instead of the hash
constexpr int kHashProbes = 8;
constexpr unsigned kNoTile = 0xFFFFFFFFu;
constexpr unsigned kKeyEmpty = 0xFFFFFFFFu;  // touches no tile
constexpr unsigned kKeyWide = 0xFFFFFFFEu;   // more than kMaxTiles tiles: the finish pass
constexpr int kSBigVox = 256;         // surface: tile-clipped boxes above this (or wider than 32 in x) are shared
constexpr int kSRecs = 32;            // shared surface triangles whose records are in LDS at once
constexpr int kSBigCols = 48;         // solid: tile-clipped column counts above this are shared
constexpr int kSQueue = 64;           // solid: shared triangles per round
constexpr int kFinishBlocks = 64;

// misc words of the workspace (zero between calls)
enum : int { SM_UNITS = 0, SM_WIDE = 1, SM_SPILL = 2, SM_SPILL_OVF = 3, SM_FINISH_DONE = 4 };

// key: ty0 (bits 0-12) | tz0 (13-25) | dy (26-28) | dz (29-31), dy/dz = extent - 1
__device__ __forceinline__ unsigned pack_key(int ty0, int tz0, int dy, int dz) {
    return (unsigned)ty0 | ((unsigned)tz0 << 13) | ((unsigned)dy << 26) | ((unsigned)dz << 29);
}
__device__ __forceinline__ void unpack_key(unsigned k, int& ty0, int& tz0, int& dy, int& dz) {
    ty0 = (int)(k & 0x1FFFu);
    tz0 = (int)((k >> 13) & 0x1FFFu);
    dy = (int)((k >> 26) & 7u);
    dz = (int)(k >> 29);
}

// Column range of a valid solid triangle clamped to the grid (the reference
// loops from its unclamped bounds; cpu_voxelizer.cpp:288-290), with the
// per-triangle out-of-grid events the oracle counts.  False when empty.
__device__ __forceinline__ bool solid_range(const SolidTri& S, int G, int& ylo, int& yhi, int& zlo, int& zhi,
                                            unsigned long long& events) {
    ylo = (int)S.ymin;
    zlo = (int)S.zmin;
    if (ylo < 0) { events++; ylo = 0; }
    if (zlo < 0) { events++; zlo = 0; }
    if (S.ymax > (float)(G - 1)) events++;
    if (S.zmax > (float)(G - 1)) events++;
    // (float)y <= ymax for an integer-valued ymax <=> y <= ymax
    yhi = S.ymax >= (float)(G - 1) ? G - 1 : (S.ymax >= 0.0f ? (int)S.ymax : -1);
    zhi = S.zmax >= (float)(G - 1) ? G - 1 : (S.zmax >= 0.0f ? (int)S.zmax : -1);
    return ylo <= yhi && zlo <= zhi;
}

// Tile -> count, merged in LDS; a tile that finds no slot within
// kHashProbes goes straight to `fallback` (global).  Returns the slot or -1.
__device__ __forceinline__ int hash_slot(unsigned* hk, unsigned tile) {
    unsigned h = (tile * 2654435761u) >> (32 - 9);
#pragma unroll 1
    for (int p = 0; p < kHashProbes; p++) {
        unsigned k = hk[h];
        if (k == kNoTile) {
            k = atomicCAS(&hk[h], kNoTile, tile);
            if (k == kNoTile) k = tile;
        }
        if (k == tile) return (int)h;
        h = (h + 1) & (kHash - 1);
    }
    return -1;
}

// Runs of consecutive active lanes with equal keys (a spatially coherent mesh
// gives neighbouring lanes the same tile): per lane, whether it heads its run,
// the run's length (valid at the head), its rank in the run and the head's
// lane.  Every lane of the wave must call it (uniform control flow).
struct WaveRun {
    bool head;
    unsigned len, rank;
    int head_lane;
};

__device__ __forceinline__ WaveRun wave_run(unsigned key, bool active) {
    const int lane = __lane_id();
    const unsigned long long act = __ballot(active);
    const unsigned prev_key = __shfl_up(key, 1);
    const bool prev_act = lane > 0 && ((act >> (lane - 1)) & 1ull);
    WaveRun r;
    r.head = active && !(prev_act && prev_key == key);
    const unsigned long long heads = __ballot(r.head);
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const unsigned long long hb = heads & upto;
    r.head_lane = hb ? 63 - __clzll((long long)hb) : lane;
    const unsigned long long breaks = (heads | ~act) & ~upto;  // the first break above me ends my run
    const int end = breaks ? __ffsll((long long)breaks) - 1 : 64;
    r.len = (unsigned)(end - lane);
    r.rank = (unsigned)(lane - r.head_lane);
    return r;
}

// Slot of a tile already in the hash (lookup along the insertion's probe
// sequence), or -1.
__device__ __forceinline__ int hash_find(const unsigned* hk, unsigned tile) {
    unsigned h = (tile * 2654435761u) >> (32 - 9);
#pragma unroll 1
    for (int p = 0; p < kHashProbes; p++) {
        const unsigned k = hk[h];
        if (k == tile) return (int)h;
        if (k == kNoTile) return -1;
        h = (h + 1) & (kHash - 1);
    }
    return -1;
}

// ---------------------------------------------------------------------------
// 1: counts, keys, wide list
// ---------------------------------------------------------------------------
// Work unit: kBinChunk consecutive triangles (kBinIters rounds of one per
// thread, staged through LDS with the next round's loads in flight).  The
// workgroup merges its (tile, count) pairs in an LDS hash, then reserves one
// range per touched tile with a single returning atomicAdd on that tile's
// count and records (tile, base) for the scatter pass, which so needs no
// global atomics.  Pairs that find no hash slot are counted apart (cntf) and
// placed by the scatter pass with global atomics.
template <bool SOLID, bool DIRECT>
__global__ void __launch_bounds__(kSBlock) slab_count_kernel(vox_info info, const float* __restrict__ tris,
                                                             SlabPlan P, SlabWs ws, unsigned long long* offrow) {
    __shared__ float s_tri[kSBlock * 9];
    __shared__ unsigned hk[DIRECT ? 1 : kHash], hv[DIRECT ? kDirectTiles : kHash];
    __shared__ unsigned s_nrec;
    const uint64_t n = info.n_triangles;
    const uint64_t first = (uint64_t)blockIdx.x * kBinChunk;
    if (DIRECT) {
        for (int i = threadIdx.x; i < P.ntiles; i += kSBlock) hv[i] = 0u;
    } else {
        for (int i = threadIdx.x; i < kHash; i += kSBlock) {
            hk[i] = kNoTile;
            hv[i] = 0u;
        }
    }
    if (threadIdx.x == 0) s_nrec = 0u;
    const int G = P.G;
    unsigned long long events = 0;
    float pre[9];
    auto fetch = [&](uint64_t c0) {  // this thread's share of the 9 * kSBlock floats from triangle c0 on
        const uint64_t lim = c0 < n ? (n - c0) * 9 : 0;
        const float* src = tris + c0 * 9;
#pragma unroll
        for (int u = 0; u < 9; u++) {
            const uint64_t i = (uint64_t)threadIdx.x + (uint64_t)u * kSBlock;
            pre[u] = i < lim ? src[i] : 0.0f;
        }
    };
    fetch(first);
    for (int it = 0; it < kBinIters; it++) {
        const uint64_t c0 = first + (uint64_t)it * kSBlock;
        if (c0 >= n) break;  // uniform over the workgroup
        __syncthreads();     // the previous iteration has read s_tri
#pragma unroll
        for (int u = 0; u < 9; u++) s_tri[threadIdx.x + u * kSBlock] = pre[u];
        if (it + 1 < kBinIters) fetch(c0 + kSBlock);  // in flight during this round
        __syncthreads();
        const uint64_t t = c0 + threadIdx.x;
        int ty0 = 0, tz0 = 0, dy = -1, dz = -1;
        if (t < n) {
            float v[3][3];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                v[k][0] = s_tri[threadIdx.x * 9 + 3 * k + 0] - info.bbox_min[0];
                v[k][1] = s_tri[threadIdx.x * 9 + 3 * k + 1] - info.bbox_min[1];
                v[k][2] = s_tri[threadIdx.x * 9 + 3 * k + 2] - info.bbox_min[2];
            }
            int ylo, yhi, zlo, zhi;
            bool any;
            if (SOLID) {
                SolidTri S;
                solid_setup(v, info, S);
                unsigned long long ev = 0;
                any = S.valid && solid_range(S, G, ylo, yhi, zlo, zhi, ev);
                if (S.valid) events += ev;
            } else {
                int lo[3], hi[3];
                surface_box(v, info, lo, hi);
                ylo = lo[1]; yhi = hi[1]; zlo = lo[2]; zhi = hi[2];
                any = true;
            }
            unsigned key = kKeyEmpty;
            if (any) {
                ty0 = ylo / P.ty;
                tz0 = zlo / P.tz;
                dy = yhi / P.ty - ty0;
                dz = zhi / P.tz - tz0;
                if ((dy + 1) * (dz + 1) > kMaxTiles) {
                    key = kKeyWide;
                    ws.wide[atomicAdd(&ws.misc[SM_WIDE], 1u)] = (unsigned)t;
                    dy = dz = -1;
                } else {
                    key = pack_key(ty0, tz0, dy, dz);
                }
            }
            ws.key[t] = key;
        }
        // the k-th tile of every lane's range, one wave-wide step per k; a run
        // of lanes with the same tile adds its length once
        const int nt = (dy + 1) * (dz + 1);
        int wmax = nt;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) wmax = max(wmax, __shfl_xor(wmax, d));
        for (int k = 0; k < wmax; k++) {
            const bool act = k < nt;
            const unsigned tile =
                act ? (unsigned)(ty0 + k % (dy + 1)) + (unsigned)(tz0 + k / (dy + 1)) * (unsigned)P.nty : 0u;
            const WaveRun r = wave_run(tile, act);
            if (r.head) {
                if (DIRECT) {
                    atomicAdd(&hv[tile], r.len);
                } else {
                    const int s = hash_slot(hk, tile);
                    if (s >= 0)
                        atomicAdd(&hv[s], r.len);
                    else
                        atomicAdd(&ws.cntf[tile], r.len);
                }
            }
        }
    }
    if (SOLID && offrow && events) atomicAdd(offrow, events);
    __syncthreads();
    // one range per touched tile; (tile, base within the tile's hashed part, slot) recorded compactly
    unsigned* rec = ws.brec + (uint64_t)blockIdx.x * ws.brec_stride;
    const int nslots = DIRECT ? P.ntiles : kHash;
    for (int i = threadIdx.x; i < nslots; i += kSBlock) {
        const unsigned tile = DIRECT ? (unsigned)i : hk[i];
        if (tile != kNoTile && hv[i] != 0u) {
            const unsigned base = atomicAdd(&ws.cnt[tile], hv[i]);
            const unsigned j = atomicAdd(&s_nrec, 1u);
            rec[3 * j] = tile;
            rec[3 * j + 1] = base;
            rec[3 * j + 2] = (unsigned)i;  // its slot: the scatter pass rebuilds the same table
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) ws.brec_n[blockIdx.x] = s_nrec;
}

// ---------------------------------------------------------------------------
// 1b: one workgroup scans the counts into list offsets (and returns the counts
// to zero) and lists the raster pass's work units: a surface tile with more
// than `split` entries is cut along x into p = 2, 4 or 8 sub-tiles (rows of
// at least 4 words), each a unit of its own -- a tile holds thousands of
// entries only where a surface runs along x through it, so its triangles
// spread over the sub-tiles.  unit = tile << 8 | log2(p) << 4 | sub.
// ---------------------------------------------------------------------------
constexpr int kScanThreads = 1024;

__global__ void __launch_bounds__(kScanThreads) slab_scan_kernel(SlabPlan P, SlabWs ws, unsigned split,
                                                                 int max_lgp) {
    __shared__ unsigned s_e[kScanThreads], s_u[kScanThreads];
    const int nt = P.ntiles;
    const int per = (nt + kScanThreads - 1) / kScanThreads;
    const int b0 = (int)threadIdx.x * per;
    const int b1 = b0 + per < nt ? b0 + per : nt;
    unsigned le = 0, lu = 0;
    for (int i = b0; i < b1; i++) {
        const unsigned c = ws.cnt[i] + ws.cntf[i];
        le += c;
        int lg = 0;
        while (lg < max_lgp && c > (split << lg)) lg++;
        lu += 1u << lg;
    }
    s_e[threadIdx.x] = le;
    s_u[threadIdx.x] = lu;
    __syncthreads();
    // exclusive scans of both partial sums (Hillis-Steele over 1024 entries)
    for (int d = 1; d < kScanThreads; d <<= 1) {
        const unsigned ae = threadIdx.x >= (unsigned)d ? s_e[threadIdx.x - d] : 0u;
        const unsigned au = threadIdx.x >= (unsigned)d ? s_u[threadIdx.x - d] : 0u;
        __syncthreads();
        s_e[threadIdx.x] += ae;
        s_u[threadIdx.x] += au;
        __syncthreads();
    }
    unsigned e = s_e[threadIdx.x] - le, u = s_u[threadIdx.x] - lu;
    for (int i = b0; i < b1; i++) {
        const unsigned ch = ws.cnt[i], c = ch + ws.cntf[i];
        ws.off[i] = e;
        ws.cur[i] = e + ch;  // unhashed pairs go after the hashed reservations
        ws.cnt[i] = 0u;
        ws.cntf[i] = 0u;
        e += c;
        int lg = 0;
        while (lg < max_lgp && c > (split << lg)) lg++;
        for (int sub = 0; sub < (1 << lg); sub++) ws.units[u + sub] = ((unsigned)i << 8) | ((unsigned)lg << 4) | (unsigned)sub;
        u += 1u << lg;
    }
    if (threadIdx.x == kScanThreads - 1) {
        ws.off[nt] = e;
        ws.misc[SM_UNITS] = u;
    }
}

// ---------------------------------------------------------------------------
// 2: per-tile lists
// ---------------------------------------------------------------------------
// Same work units as the count pass.  The workgroup loads its recorded (tile,
// base) pairs into an LDS hash; a triangle's slot in tile t is off[t] + base
// + the rank the LDS counter gives it.  Tiles the count pass could not hash
// take slots past the hashed part with global atomics (cur, set by the scan).
template <bool DIRECT>
__global__ void __launch_bounds__(kSBlock) slab_scatter_kernel(uint64_t n, SlabPlan P, SlabWs ws) {
    // hashed: hk = tile, hv = list position of the block's range, hc = slots
    // taken; direct: hv[tile] = next list position of the block's range
    __shared__ unsigned hk[DIRECT ? 1 : kHash], hv[DIRECT ? kDirectTiles : kHash], hc[DIRECT ? 1 : kHash];
    if (!DIRECT) {
        for (int i = threadIdx.x; i < kHash; i += kSBlock) {
            hk[i] = kNoTile;
            hc[i] = 0u;
        }
        __syncthreads();
    }
    // the count pass's table, slot for slot, so lookups follow its probe sequences
    const unsigned nrec = ws.brec_n[blockIdx.x];
    const unsigned* rec = ws.brec + (uint64_t)blockIdx.x * ws.brec_stride;
    for (unsigned j = threadIdx.x; j < nrec; j += kSBlock) {
        const unsigned tile = rec[3 * j];
        const unsigned slot = DIRECT ? tile : rec[3 * j + 2] & (kHash - 1);
        if (!DIRECT) hk[slot] = tile;
        hv[slot] = ws.off[tile] + rec[3 * j + 1];
    }
    __syncthreads();
    const uint64_t first = (uint64_t)blockIdx.x * kBinChunk;
    for (int it = 0; it < kBinIters; it++) {
        const uint64_t t = first + (uint64_t)it * kSBlock + threadIdx.x;
        if (first + (uint64_t)it * kSBlock >= n) break;  // uniform over the workgroup
        const unsigned key = t < n ? ws.key[t] : kKeyEmpty;
        int ty0 = 0, tz0 = 0, dy = -1, dz = -1;
        if (key != kKeyEmpty && key != kKeyWide) unpack_key(key, ty0, tz0, dy, dz);
        const int nt = (dy + 1) * (dz + 1);
        int wmax = nt;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) wmax = max(wmax, __shfl_xor(wmax, d));
        for (int k = 0; k < wmax; k++) {
            const bool act = k < nt;
            const unsigned tile =
                act ? (unsigned)(ty0 + k % (dy + 1)) + (unsigned)(tz0 + k / (dy + 1)) * (unsigned)P.nty : 0u;
            const WaveRun r = wave_run(tile, act);
            unsigned pos = 0;
            if (r.head) {
                if (DIRECT) {
                    pos = atomicAdd(&hv[tile], r.len);
                } else {
                    const int s = hash_find(hk, tile);
                    pos = s >= 0 ? hv[s] + atomicAdd(&hc[s], r.len) : atomicAdd(&ws.cur[tile], r.len);
                }
            }
            pos = __shfl(pos, r.head_lane) + r.rank;
            if (act) ws.list[pos] = (unsigned)t;
        }
    }
}

// ---------------------------------------------------------------------------
// 3a: surface raster
// ---------------------------------------------------------------------------
// A lane's own triangle (x extent <= 32), its box clipped to the tile in y/z:
// the ZX edge tests once per z slice into a mask over x, then per row the YZ
// test and, at the x that passed ZX, the plane and XY tests; the passing bits
// are ORed into the tile's LDS row.
__device__ __forceinline__ void slab_lane_raster(uint32_t* lds, const SurfTri& T, const vox_info& info, int xa,
                                                 int xb, int ya, int yb, int za, int zb, int y0, int z0, int xu0,
                                                 int ty, int rw) {
    const float ux = info.unit[0], uy = info.unit[1], uz = info.unit[2];
    const int c0 = xa & ~31, sh = xa - c0;
    for (int z = za; z <= zb; z++) {
        const float pz = (float)z * uz;
        float azx[3];
#pragma unroll
        for (int i = 0; i < 3; i++) azx[i] = T.zx.a[i] * pz;
        uint32_t zxm = 0;  // bit i: x = xa + i passes the ZX edges
        for (int x = xa; x <= xb; x++) {
            const float px = (float)x * ux;
            bool ok = true;
#pragma unroll
            for (int i = 0; i < 3; i++) {
                float b = azx[i];
                b = b + T.zx.b[i] * px;
                b = b + T.zx.d[i];
                ok = ok && !(b < 0.0f);
            }
            if (ok) zxm |= 1u << (x - xa);
        }
        if (!zxm) continue;
        const float nzpz = T.n[2] * pz;
        const int rowz = (z - z0) * ty - y0;
        for (int y = ya; y <= yb; y++) {
            const float py = (float)y * uy;
            if (!edges_pass(T.yz, py, pz)) continue;
            const float nypy = T.n[1] * py;
            float bxy[3];
#pragma unroll
            for (int i = 0; i < 3; i++) bxy[i] = T.xy.b[i] * py;
            uint32_t m = 0, cand = zxm;
            while (cand) {
                const int k = __ffs(cand) - 1;
                cand &= cand - 1;
                const float px = (float)(xa + k) * ux;
                float t = T.n[0] * px;
                t = t + nypy;
                t = t + nzpz;
                const float r = (t + T.d1) * (t + T.d2);
                bool ok = !(r > 0.0f);
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    float a = T.xy.a[i] * px;
                    a = a + bxy[i];
                    a = a + T.xy.d[i];
                    ok = ok && !(a < 0.0f);
                }
                if (ok) m |= 1u << k;
            }
            if (!m) continue;
            const uint64_t M = (uint64_t)m << sh;  // bit k = voxel c0 + k
            const int w = (rowz + y) * rw + ((c0 - xu0) >> 5);
            if ((uint32_t)M) atomicOr(lds + w, __brev((uint32_t)M));
            if (M >> 32) atomicOr(lds + w + 1, __brev((uint32_t)(M >> 32)));
        }
    }
}

// One work unit of the raster pass: a tile, or one of its 2^lgp sub-tiles
// along x (x0 .. x0 + G/2^lgp - 1, rws = rw >> lgp words per row).
struct SlabTile {
    int tile;
    int y0, z0;     // first row of the tile
    int x0, x1;     // x range of the unit
    int rws;        // words per row of the unit
    int wofs;       // first word of the unit inside a table row
};

__device__ __forceinline__ SlabTile slab_unit(const SlabPlan& P, unsigned unit) {
    SlabTile S;
    S.tile = (int)(unit >> 8);
    const int lgp = (int)((unit >> 4) & 15u), sub = (int)(unit & 15u);
    S.y0 = (S.tile % P.nty) * P.ty;
    S.z0 = (S.tile / P.nty) * P.tz;
    S.rws = P.rw >> lgp;
    S.wofs = sub * S.rws;
    S.x0 = S.wofs * 32;
    S.x1 = S.x0 + S.rws * 32 - 1;
    return S;
}

// Zero the LDS tile (16-byte stores).
template <int NT>
__device__ __forceinline__ void slab_clear(uint32_t* lds, int words) {
    uint4* l4 = (uint4*)lds;
    for (int i = threadIdx.x; i < words / 4; i += NT) l4[i] = make_uint4(0u, 0u, 0u, 0u);
}

// Write the LDS unit to the table with 16-byte stores.  A whole-row unit is
// tz segments of ty*rw contiguous words (rows y0..y0+ty-1 of one z; ty*rw is
// a multiple of 4); a sub-tile writes rws words (at least 4) at word wofs of
// each of its rows.  acc: combine with the table (OR for surface, XOR for
// solid).
template <bool XOR, int NT>
__device__ __forceinline__ void slab_flush(const SlabPlan& P, const SlabTile& S, const uint32_t* lds,
                                           uint32_t* __restrict__ table, bool acc) {
    const bool whole = S.rws == P.rw;
    const int seg4 = whole ? (P.ty * P.rw) >> 2 : S.rws >> 2;  // uint4 per contiguous segment
    const int nseg = whole ? P.tz : P.ty * P.tz;
    const int total4 = seg4 * nseg;
    const uint4* l4 = (const uint4*)lds;
    for (int i = threadIdx.x; i < total4; i += NT) {
        const int g = i / seg4, k = i - g * seg4;
        const int y = whole ? S.y0 : S.y0 + g % P.ty, z = whole ? S.z0 + g : S.z0 + g / P.ty;
        uint4* dst = (uint4*)(table + ((uint64_t)z * (uint64_t)P.G + (uint64_t)y) * (uint64_t)P.rw +
                              (uint64_t)S.wofs) + k;
        uint4 v = l4[i];
        if (acc) {
            const uint4 o = *dst;
            if (XOR) {
                v.x ^= o.x; v.y ^= o.y; v.z ^= o.z; v.w ^= o.w;
            } else {
                v.x |= o.x; v.y |= o.y; v.z |= o.z; v.w |= o.w;
            }
        }
        *dst = v;
    }
}

// Work units are taken in a grid-stride loop (their number is known on the
// device only, after the scan).
template <int NT>
__global__ void __launch_bounds__(NT, 4) slab_surface_kernel(vox_info info, const float* __restrict__ tris,
                                                                 uint32_t* __restrict__ table, SlabPlan P,
                                                                 SlabWs ws, int accumulate) {
    extern __shared__ uint4 lds4[];
    uint32_t* lds = (uint32_t*)lds4;
    __shared__ SurfRec recs[kSRecs];
    __shared__ unsigned long long incl[kSRecs];
    __shared__ unsigned wq[NT / 64];
    const unsigned nunits = ws.misc[SM_UNITS];
    const float ux = info.unit[0], uy = info.unit[1], uz = info.unit[2];
    const int wave = (int)(threadIdx.x >> 6);
    for (unsigned unit_i = blockIdx.x; unit_i < nunits; unit_i += gridDim.x) {
        const SlabTile St = slab_unit(P, ws.units[unit_i]);
        const int words = P.ty * P.tz * St.rws;
        __syncthreads();  // the previous unit's flush has read the LDS
        slab_clear<NT>(lds, words);
        const unsigned beg = ws.off[St.tile], end = ws.off[St.tile + 1];
        __syncthreads();
        const int ylim = St.y0 + P.ty - 1, zlim = St.z0 + P.tz - 1;
        for (unsigned base = beg; base < end; base += NT) {
            const unsigned i = base + threadIdx.x;
            const bool have = i < end;
            const unsigned id = have ? ws.list[i] : 0u;
            float v[3][3];
            SurfTri T;
            bool big = false;
            if (have) {
                load_tri(tris, id, info, v);
                surface_box(v, info, T.lo, T.hi);
                T.lo[0] = max(T.lo[0], St.x0); T.hi[0] = min(T.hi[0], St.x1);
                T.lo[1] = max(T.lo[1], St.y0); T.hi[1] = min(T.hi[1], ylim);
                T.lo[2] = max(T.lo[2], St.z0); T.hi[2] = min(T.hi[2], zlim);
                const int nx = T.hi[0] - T.lo[0] + 1;
                if (nx > 0) {  // the triangle reaches this unit's x range
                    const int64_t vol = (int64_t)nx * (T.hi[1] - T.lo[1] + 1) * (T.hi[2] - T.lo[2] + 1);
                    big = vol > kSBigVox || nx > 32;
                    if (!big) {
                        surface_setup_tests(v, info, T);
                        slab_lane_raster(lds, T, info, T.lo[0], T.hi[0], T.lo[1], T.hi[1], T.lo[2], T.hi[2],
                                         St.y0, St.z0, St.x0, P.ty, St.rws);
                    }
                }
            }
            // large boxes: their (row, 32-voxel x chunk) items are dealt over the workgroup
            const unsigned long long bal = __ballot(big);
            if (__lane_id() == 0) wq[wave] = (unsigned)__popcll(bal);
            __syncthreads();
            int myq = -1, nq = 0;
#pragma unroll
            for (int w = 0; w < NT / 64; w++) {
                if (w == wave) myq = nq + (int)__popcll(bal & ((1ull << __lane_id()) - 1ull));
                nq += (int)wq[w];
            }
            if (!big) myq = -1;
            for (int r0 = 0; r0 < nq; r0 += kSRecs) {
                const int nr = nq - r0 < kSRecs ? nq - r0 : kSRecs;
                if (myq >= r0 && myq < r0 + nr) {
                    surface_setup_tests(v, info, T);
                    const int xa = T.lo[0], xb = T.hi[0], ya = T.lo[1], yb = T.hi[1], za = T.lo[2], zb = T.hi[2];
                    int nk = 1, band = 0;
                    float sx = 0.0f;
                    band_chunks(T, info, xa, xb, yb, zb, nk, band, sx);
                    publish(recs[myq - r0], T, xa, xb, ya, yb, za, zb, nk, band, sx);
                    incl[myq - r0] = (unsigned long long)(yb - ya + 1) * (unsigned long long)(zb - za + 1) *
                                     (unsigned long long)nk;
                }
                __syncthreads();
                if (threadIdx.x < 64) {  // inclusive scan of the round's counts (nr <= 64)
                    const int lane = (int)threadIdx.x;
                    unsigned long long c = lane < nr ? incl[lane] : 0ull;
#pragma unroll
                    for (int d = 1; d < 64; d <<= 1) {
                        const unsigned long long o = __shfl_up(c, d);
                        if (lane >= d) c += o;
                    }
                    if (lane < nr) incl[lane] = c;
                }
                __syncthreads();
                const unsigned long long total = incl[nr - 1];
                for (unsigned long long j = threadIdx.x; j < total; j += NT) {
                    int lo = 0, hi = nr - 1;
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (incl[mid] > j) hi = mid; else lo = mid + 1;
                    }
                    const SurfRec& R = recs[lo];
                    const unsigned long long q = j - (lo ? incl[lo - 1] : 0ull);
                    const unsigned nk = (unsigned)R.nk, ny = (unsigned)(R.yb - R.ya + 1);
                    const unsigned row = (unsigned)(q / nk);
                    const int k = (int)(q - (unsigned long long)row * nk);
                    const int y = R.ya + (int)(row % ny), z = R.za + (int)(row / ny);
                    const float py = (float)y * uy, pz = (float)z * uz;
                    if (!pass3(R.yz, py, pz)) continue;
                    int bx0 = R.xa, bx1 = R.xb;
                    if (R.band) {
                        plane_band(R, ux, py, pz, bx0, bx1);
                        if (bx0 > bx1) continue;
                    }
                    const int cx = ((bx0 >> 5) + k) << 5;
                    if (cx > bx1) continue;
                    const int x0 = cx > bx0 ? cx : bx0, x1 = (cx + 31) < bx1 ? (cx + 31) : bx1;
                    const uint32_t m = row_mask(R.n, R.d1, R.d2, R.xy, R.zx, ux, py, pz, x0, x1, cx);
                    if (m)
                        atomicOr(lds + ((y - St.y0) + (z - St.z0) * P.ty) * St.rws + ((cx - St.x0) >> 5), __brev(m));
                }
                __syncthreads();  // the records are rewritten by the next round
            }
        }
        __syncthreads();
        slab_flush<false, NT>(P, St, lds, table, accumulate != 0);
    }
}

// ---------------------------------------------------------------------------
// 3b: solid raster
// ---------------------------------------------------------------------------
// One hit column: the marker at its run end (x = G-1 when the run leaves the
// row, whose remainder is recorded for the finish pass), and its events.
__device__ __forceinline__ void slab_column(uint32_t* lds, int lrow, int rw, int64_t xmax, int y, int z, int G,
                                            uint64_t total_bits, const SlabWs& ws, unsigned long long& events) {
    if (xmax < 0 || xmax >= (int64_t)G) events++;
    if (xmax < 0) return;
    const uint32_t xm = xmax < (int64_t)G ? (uint32_t)xmax : (uint32_t)(G - 1);
    atomicXor(lds + lrow * rw + (int)(xm >> 5), 1u << (31u - (xm & 31u)));
    if (xmax >= (int64_t)G) {
        const uint64_t row = (uint64_t)y * (uint64_t)G + (uint64_t)z * (uint64_t)G * (uint64_t)G;
        if (row + (uint64_t)G < total_bits) {
            const uint64_t last = (uint64_t)xmax < total_bits - 1 - row ? row + (uint64_t)xmax : total_bits - 1;
            const unsigned slot = atomicAdd(&ws.misc[SM_SPILL], 1u);
            if (slot < kSlabSpillCap) {
                ws.spill[2 * slot] = row + (uint64_t)G;
                ws.spill[2 * slot + 1] = last;
            } else {
                ws.misc[SM_SPILL_OVF] = 1u;
            }
        }
    }
}

// Suffix XOR of every row of the LDS tile in place: bit x becomes the XOR of
// the markers at x' >= x.  Bits are MSB-first (x = 32w + j at bit 31-j), so
// within a word that is a prefix XOR from bit 0 upward; the words of larger x
// contribute their marker parity (a ballot over the row's lanes).  rw <= 64.
template <int NT>
__device__ __forceinline__ void slab_suffix(uint32_t* lds, int rows, int rw) {
    const int per_wave = 64 / rw;
    const int lane = __lane_id();
    const int wave = (int)(threadIdx.x >> 6), nwaves = NT / 64;
    const int my_r = lane / rw, my_w = lane - my_r * rw;
    for (int r0 = wave * per_wave; r0 < rows; r0 += nwaves * per_wave) {
        const int r = r0 + my_r;
        const bool act = my_r < per_wave && r < rows;
        uint32_t m = act ? lds[r * rw + my_w] : 0u;
        const unsigned long long par = __ballot(act && (__popc(m) & 1));
        uint32_t s = m;
        s ^= s << 1;
        s ^= s << 2;
        s ^= s << 4;
        s ^= s << 8;
        s ^= s << 16;
        const int hi_cnt = rw - 1 - my_w;  // words of my row with a larger x: lanes my_w+1 .. rw-1
        unsigned long long hi_mask = 0ull;
        if (act && hi_cnt > 0) hi_mask = ((1ull << hi_cnt) - 1ull) << (my_r * rw + my_w + 1);
        const uint32_t carry = (uint32_t)(__popcll(par & hi_mask) & 1);
        if (act) lds[r * rw + my_w] = s ^ (carry ? 0xFFFFFFFFu : 0u);
    }
}

template <int NT>
__global__ void __launch_bounds__(NT, 4) slab_solid_kernel(vox_info info, const float* __restrict__ tris,
                                                               uint32_t* __restrict__ table, SlabPlan P, SlabWs ws,
                                                               int accumulate, unsigned long long* offrow) {
    extern __shared__ uint4 lds4[];
    uint32_t* lds = (uint32_t*)lds4;
    __shared__ SolidTri sq[kSQueue];
    __shared__ int sqr[kSQueue][4];  // tile-clipped column range of a shared triangle
    __shared__ unsigned sqn;
    const int G = P.G;
    const uint64_t total_bits = (uint64_t)G * (uint64_t)G * (uint64_t)G;
    const unsigned nunits = ws.misc[SM_UNITS];  // whole tiles: solid tiles are never split
    unsigned long long events = 0;
    for (unsigned unit_i = blockIdx.x; unit_i < nunits; unit_i += gridDim.x) {
        const SlabTile St = slab_unit(P, ws.units[unit_i]);
        __syncthreads();  // the previous unit's flush has read the LDS
        slab_clear<NT>(lds, P.tile_words);
        const unsigned beg = ws.off[St.tile], end = ws.off[St.tile + 1];
        if (threadIdx.x == 0) sqn = 0u;
        __syncthreads();
        const int ylim = St.y0 + P.ty - 1, zlim = St.z0 + P.tz - 1;
        for (unsigned base = beg; base < end; base += NT) {
            const unsigned i = base + threadIdx.x;
            if (i < end) {
                float v[3][3];
                load_tri(tris, ws.list[i], info, v);
                SolidTri S;
                solid_setup(v, info, S);
                int ylo, yhi, zlo, zhi;
                unsigned long long dummy = 0;
                solid_range(S, G, ylo, yhi, zlo, zhi, dummy);  // binned triangles are valid and non-empty
                ylo = max(ylo, St.y0); yhi = min(yhi, ylim);
                zlo = max(zlo, St.z0); zhi = min(zhi, zlim);
                const int cols = (yhi >= ylo && zhi >= zlo) ? (yhi - ylo + 1) * (zhi - zlo + 1) : 0;
                bool shared = false;
                if (cols > kSBigCols) {
                    const unsigned q = atomicAdd(&sqn, 1u);
                    if (q < (unsigned)kSQueue) {
                        sq[q] = S;
                        sqr[q][0] = ylo; sqr[q][1] = yhi; sqr[q][2] = zlo; sqr[q][3] = zhi;
                        shared = true;
                    }
                }
                if (!shared) {
                    for (int y = ylo; y <= yhi; y++) {
                        const float py = ((float)y + 0.5f) * info.unit[1];
                        for (int z = zlo; z <= zhi; z++) {
                            const float pz = ((float)z + 0.5f) * info.unit[2];
                            if (!column_hit(S, py, pz)) continue;
                            slab_column(lds, (y - St.y0) + (z - St.z0) * P.ty, P.rw, column_xmax(S, info, py, pz), y, z,
                                        G, total_bits, ws, events);
                        }
                    }
                }
            }
            __syncthreads();
            const unsigned nq = sqn < (unsigned)kSQueue ? sqn : (unsigned)kSQueue;
            __syncthreads();  // every thread has read sqn
            if (threadIdx.x == 0) sqn = 0u;  // the next round's first atomicAdd follows the barrier below
            for (unsigned q = 0; q < nq; q++) {
                const SolidTri& S = sq[q];
                const int ylo = sqr[q][0], yhi = sqr[q][1], zlo = sqr[q][2], zhi = sqr[q][3];
                const int nz = zhi - zlo + 1;
                const int cols = (yhi - ylo + 1) * nz;
                for (int c = threadIdx.x; c < cols; c += NT) {
                    const int y = ylo + c / nz, z = zlo + c % nz;
                    const float py = ((float)y + 0.5f) * info.unit[1];
                    const float pz = ((float)z + 0.5f) * info.unit[2];
                    if (!column_hit(S, py, pz)) continue;
                    slab_column(lds, (y - St.y0) + (z - St.z0) * P.ty, P.rw, column_xmax(S, info, py, pz), y, z, G,
                                total_bits, ws, events);
                }
            }
            __syncthreads();  // the queue entries are rewritten by the next round
        }
        __syncthreads();
        slab_suffix<NT>(lds, P.ty * P.tz, P.rw);
        __syncthreads();
        slab_flush<true, NT>(P, St, lds, table, accumulate != 0);
    }
    if (offrow && events) atomicAdd(offrow, events);
}

// ---------------------------------------------------------------------------
// 4: wide triangles, spilled runs, counter reset
// ---------------------------------------------------------------------------
template <bool SOLID>
__global__ void __launch_bounds__(kSBlock) slab_finish_kernel(vox_info info, const float* __restrict__ tris,
                                                              uint32_t* __restrict__ table, SlabPlan P, SlabWs ws,
                                                              unsigned long long* offrow) {
    const unsigned nw = ws.misc[SM_WIDE];
    const int G = P.G;
    const uint64_t total_bits = (uint64_t)G * (uint64_t)G * (uint64_t)G;
    unsigned long long events = 0;
    for (unsigned w = blockIdx.x; w < nw; w += gridDim.x) {
        const unsigned id = ws.wide[w];
        if ((uint64_t)id >= info.n_triangles) continue;  // written by the count pass, always < n
        float v[3][3];
        load_tri(tris, id, info, v);
        if (!SOLID) {
            SurfTri T;
            surface_setup(v, info, T);
            const int ny = T.hi[1] - T.lo[1] + 1, nz = T.hi[2] - T.lo[2] + 1;
            const int c0 = T.lo[0] >> 5, nch = (T.hi[0] >> 5) - c0 + 1;
            const uint64_t items = (uint64_t)ny * nz * nch;
            for (uint64_t k = threadIdx.x; k < items; k += kSBlock) {
                const int ch = (int)(k % nch);
                const uint64_t r = k / nch;
                const int y = T.lo[1] + (int)(r % ny), z = T.lo[2] + (int)(r / ny);
                const float py = (float)y * info.unit[1], pz = (float)z * info.unit[2];
                if (!edges_pass(T.yz, py, pz)) continue;
                const int cx = (c0 + ch) << 5;
                const int xa = max(T.lo[0], cx), xb = min(T.hi[0], cx + 31);
                uint32_t mask = 0;
                for (int x = xa; x <= xb; x++) {
                    const float px = (float)x * info.unit[0];
                    if (plane_pass(T, px, py, pz) && edges_pass(T.xy, px, py) && edges_pass(T.zx, pz, px))
                        mask |= 1u << (31u - (uint32_t)(x & 31));
                }
                if (mask) atomicOr(table + (((uint64_t)y + (uint64_t)z * G) * G + (uint64_t)cx) / 32, mask);
            }
        } else {
            SolidTri S;
            solid_setup(v, info, S);
            int ylo, yhi, zlo, zhi;
            unsigned long long dummy = 0;
            solid_range(S, G, ylo, yhi, zlo, zhi, dummy);  // wide triangles are valid and non-empty
            const int nz = zhi - zlo + 1;
            const uint64_t cols = (uint64_t)(yhi - ylo + 1) * (uint64_t)nz;
            for (uint64_t c = threadIdx.x; c < cols; c += kSBlock) {
                const int y = ylo + (int)(c / nz), z = zlo + (int)(c % nz);
                const float py = ((float)y + 0.5f) * info.unit[1], pz = ((float)z + 0.5f) * info.unit[2];
                if (!column_hit(S, py, pz)) continue;
                const int64_t xmax = column_xmax(S, info, py, pz);
                if (xmax < 0 || xmax >= G) events++;
                if (xmax < 0) continue;
                xor_column_run(table, VOX_LAYOUT_LINEAR, G, y, z, xmax);
            }
        }
    }
    if (SOLID) {
        const unsigned ns = ws.misc[SM_SPILL];
        if (!ws.misc[SM_SPILL_OVF]) {
            const unsigned m = ns < kSlabSpillCap ? ns : kSlabSpillCap;
            for (unsigned s = blockIdx.x; s < m; s += gridDim.x) {
                const uint64_t b0 = ws.spill[2 * s], b1 = ws.spill[2 * s + 1];
                for (uint64_t w = (b0 >> 5) + threadIdx.x; w <= (b1 >> 5); w += kSBlock) {
                    const uint32_t lo = w == (b0 >> 5) ? (uint32_t)(b0 & 31u) : 0u;
                    const uint32_t hi = w == (b1 >> 5) ? (uint32_t)(b1 & 31u) : 31u;
                    atomicXor(table + w, (0xFFFFFFFFu >> lo) & (0xFFFFFFFFu << (31u - hi)));
                }
            }
        } else {
            // more spilled runs than records: recompute them from every binned triangle
            const uint64_t n = info.n_triangles;
            for (uint64_t t = (uint64_t)blockIdx.x * kSBlock + threadIdx.x; t < n;
                 t += (uint64_t)gridDim.x * kSBlock) {
                const unsigned key = ws.key[t];
                if (key == kKeyEmpty || key == kKeyWide) continue;  // wide runs were applied whole above
                float v[3][3];
                load_tri(tris, t, info, v);
                SolidTri S;
                solid_setup(v, info, S);
                int ylo, yhi, zlo, zhi;
                unsigned long long dummy = 0;
                solid_range(S, G, ylo, yhi, zlo, zhi, dummy);
                for (int y = ylo; y <= yhi; y++) {
                    const float py = ((float)y + 0.5f) * info.unit[1];
                    for (int z = zlo; z <= zhi; z++) {
                        const float pz = ((float)z + 0.5f) * info.unit[2];
                        if (!column_hit(S, py, pz)) continue;
                        const int64_t xmax = column_xmax(S, info, py, pz);
                        if (xmax < (int64_t)G) continue;
                        const uint64_t row = (uint64_t)y * G + (uint64_t)z * G * G;
                        if (row + G >= total_bits) continue;
                        const uint64_t b1 =
                            (uint64_t)xmax < total_bits - 1 - row ? row + (uint64_t)xmax : total_bits - 1;
                        xor_linear_run(table, row + G, b1);
                    }
                }
            }
        }
    }
    if (SOLID && offrow && events) atomicAdd(offrow, events);
    __syncthreads();  // this workgroup has read the counters
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&ws.misc[SM_FINISH_DONE], 1u) == gridDim.x - 1) {
            ws.misc[SM_WIDE] = 0u;
            ws.misc[SM_SPILL] = 0u;
            ws.misc[SM_SPILL_OVF] = 0u;
            ws.misc[SM_FINISH_DONE] = 0u;
            __threadfence();
        }
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static int ilog2i(uint64_t v) {
    int r = 0;
    while ((2ull << r) <= v) r++;
    return r;
}

// Tiles of ty x tz rows: by default at most 32 KB of LDS and at least ~1024
// tiles when the grid allows; lg_ty / lg_tz >= 0 ask for 2^lg rows instead.
// split_lg >= 0: surface tiles above 2^split_lg entries are cut along x.
bool make_slab_plan(uint32_t G, SlabPlan& P, int lg_ty, int lg_tz, int split_lg, int wg) {
    P = SlabPlan{};
    if (G < 32 || (G & 31u) || G > 2048) return false;
    const uint64_t rw = G / 32;
    int ty, tz;
    if (lg_ty >= 0 || lg_tz >= 0) {
        ty = 1 << (lg_ty >= 0 ? lg_ty : 4);
        tz = 1 << (lg_tz >= 0 ? lg_tz : 4);
    } else {
        uint64_t rows = 8192 / rw;
        const uint64_t par = (uint64_t)G * G / 1024;
        if (par < rows) rows = par;
        if (rows < 1) rows = 1;
        rows = 1ull << ilog2i(rows);
        const int lg = ilog2i(rows);
        ty = 1 << ((lg + 1) / 2);
        tz = (int)(rows / (uint64_t)ty);
    }
    if ((uint32_t)ty > G) ty = (int)G;
    if ((uint32_t)tz > G) tz = (int)G;
    // 16-byte flush segments: ty * rw must be a multiple of 4 words
    while ((ty * (int)rw) % 4 != 0 && (uint32_t)(ty * 2) <= G) ty *= 2;
    if ((ty * (int)rw) % 4 != 0) return false;
    if (G % (uint32_t)ty != 0 || G % (uint32_t)tz != 0) return false;
    if ((uint64_t)ty * (uint64_t)tz * rw * 4 > 65536) return false;  // LDS per tile
    P.G = (int)G;
    P.ty = ty;
    P.tz = tz;
    P.nty = (int)G / ty;
    P.ntz = (int)G / tz;
    P.ntiles = P.nty * P.ntz;
    P.rw = (int)rw;
    P.tile_words = ty * tz * (int)rw;
    P.split = split_lg >= 0 ? (split_lg >= 30 ? 0x40000000u : (1u << split_lg)) : 512u;
    P.max_lgp = 0;
    while (P.max_lgp < 3 && (P.rw >> (P.max_lgp + 1)) >= 4) P.max_lgp++;
    P.wide_wg = wg == 1 ? 1 : 0;  // 1: 1024-thread raster workgroups; 2 or 0 (default): 256
    return P.nty <= 8192 && P.ntz <= 8192 && P.ntiles <= (1 << 24);
}

SlabLayout slab_layout(uint64_t tiles_cap, uint64_t n_cap) {
    SlabLayout L{};
    auto align = [](uint64_t v) { return (v + 255) & ~255ull; };
    uint64_t off = 0;
    L.misc_off = off; off = align(off + 4ull * 64);
    L.cnt_off = off; off = align(off + 4ull * (tiles_cap + 1));
    L.cntf_off = off; off = align(off + 4ull * (tiles_cap + 1));
    L.zero_bytes = off;  // misc and counts start at zero and are left zero by every call
    L.off_off = off; off = align(off + 4ull * (tiles_cap + 1));
    L.cur_off = off; off = align(off + 4ull * (tiles_cap + 1));
    L.key_off = off; off = align(off + 4ull * (n_cap > 0 ? n_cap : 1));
    L.list_off = off; off = align(off + 4ull * kMaxTiles * (n_cap > 0 ? n_cap : 1));
    L.units_off = off; off = align(off + 4ull * 8 * (tiles_cap + 1));
    const uint64_t bins = (n_cap + kBinChunk - 1) / kBinChunk + 1;
    // records per binning workgroup: at most one per hash slot, or one per tile
    // with per-tile counters (at most kDirectTiles tiles)
    // with per-tile counters (calls of at most kDirectTiles tiles; a workspace
    // sized for more tiles still serves such calls, so the stride covers
    // min(tiles_cap, kDirectTiles) records whatever plan sized it)
    const uint64_t direct_cap = tiles_cap < (uint64_t)kDirectTiles ? tiles_cap : (uint64_t)kDirectTiles;
    L.brec_stride = 3ull * (direct_cap > (uint64_t)kHash ? direct_cap : (uint64_t)kHash);
    L.brec_off = off; off = align(off + 4ull * L.brec_stride * bins);
    L.brec_n_off = off; off = align(off + 4ull * bins);
    L.wide_off = off; off = align(off + 4ull * (n_cap > 0 ? n_cap : 1));
    L.spill_off = off; off = align(off + 16ull * kSlabSpillCap);
    L.total = off;
    L.tiles_cap = tiles_cap;
    L.n_cap = n_cap;
    return L;
}

// Whether a workspace of layout L serves a call with plan P and n triangles:
// every per-call region the kernels index must fit.
bool slab_layout_fits(const SlabPlan& P, uint64_t n, const SlabLayout& L) {
    if ((uint64_t)P.ntiles > L.tiles_cap || n > L.n_cap) return false;
    const bool direct = P.ntiles <= kDirectTiles;
    const uint64_t recs = direct ? (uint64_t)P.ntiles : (uint64_t)kHash;  // records per binning workgroup
    if (3ull * recs > L.brec_stride) return false;
    const uint64_t bins = (L.n_cap + kBinChunk - 1) / kBinChunk + 1;
    if ((n + kBinChunk - 1) / kBinChunk > bins) return false;
    if (L.units_off + 4ull * 8 * ((uint64_t)P.ntiles + 1) > L.brec_off) return false;  // units fit before brec
    return true;
}

static SlabWs slab_ws(void* mem, const SlabLayout& L) {
    char* b = (char*)mem;
    SlabWs w;
    w.misc = (unsigned*)(b + L.misc_off);
    w.cnt = (unsigned*)(b + L.cnt_off);
    w.cntf = (unsigned*)(b + L.cntf_off);
    w.brec = (unsigned*)(b + L.brec_off);
    w.brec_stride = L.brec_stride;
    w.brec_n = (unsigned*)(b + L.brec_n_off);
    w.off = (unsigned*)(b + L.off_off);
    w.cur = (unsigned*)(b + L.cur_off);
    w.key = (unsigned*)(b + L.key_off);
    w.list = (unsigned*)(b + L.list_off);
    w.units = (unsigned*)(b + L.units_off);
    w.wide = (unsigned*)(b + L.wide_off);
    w.spill = (unsigned long long*)(b + L.spill_off);
    return w;
}

hipError_t launch_slab(const vox_info& info, const float* tris, uint32_t* table, const SlabPlan& P, bool solid,
                       bool accumulate, unsigned long long* offrow, void* mem, const SlabLayout& L,
                       hipStream_t stream) {
    const SlabWs ws = slab_ws(mem, L);
    const uint64_t n = info.n_triangles;
    const unsigned blocks = (unsigned)((n + kBinChunk - 1) / kBinChunk);
    const unsigned cblocks = blocks > 0 ? blocks : 1u;  // the scan runs even without triangles
    const size_t lds = (size_t)P.tile_words * 4;
    const bool direct = P.ntiles <= kDirectTiles;
    const bool wide_wg = P.wide_wg != 0;
    if (solid) {
        if (direct)
            slab_count_kernel<true, true><<<cblocks, kSBlock, 0, stream>>>(info, tris, P, ws, offrow);
        else
            slab_count_kernel<true, false><<<cblocks, kSBlock, 0, stream>>>(info, tris, P, ws, offrow);
    } else {
        if (direct)
            slab_count_kernel<false, true><<<cblocks, kSBlock, 0, stream>>>(info, tris, P, ws, nullptr);
        else
            slab_count_kernel<false, false><<<cblocks, kSBlock, 0, stream>>>(info, tris, P, ws, nullptr);
    }
    slab_scan_kernel<<<1, kScanThreads, 0, stream>>>(P, ws, solid ? 0x40000000u : P.split, solid ? 0 : P.max_lgp);
    if (blocks) {
        if (direct)
            slab_scatter_kernel<true><<<blocks, kSBlock, 0, stream>>>(n, P, ws);
        else
            slab_scatter_kernel<false><<<blocks, kSBlock, 0, stream>>>(n, P, ws);
    }
    if (solid) {
        // four workgroups of 256 threads per CU, or one of 1024 (a heavy tile gets the whole CU;
        // measured slower at 1024^3: 148 vs 109 us on the SH sphere)
        if (wide_wg)
            slab_solid_kernel<1024><<<P.ntiles, 1024, lds, stream>>>(info, tris, table, P, ws, accumulate ? 1 : 0,
                                                                    offrow);
        else
            slab_solid_kernel<256><<<P.ntiles, 256, lds, stream>>>(info, tris, table, P, ws, accumulate ? 1 : 0,
                                                                  offrow);
        slab_finish_kernel<true><<<kFinishBlocks, kSBlock, 0, stream>>>(info, tris, table, P, ws, offrow);
    } else {
        // units beyond the grid are taken by a grid-stride loop
        const unsigned grid = P.max_lgp > 0 ? (unsigned)P.ntiles * 2u : (unsigned)P.ntiles;
        if (wide_wg)
            slab_surface_kernel<1024><<<grid, 1024, lds, stream>>>(info, tris, table, P, ws, accumulate ? 1 : 0);
        else
            slab_surface_kernel<256><<<grid, 256, lds, stream>>>(info, tris, table, P, ws, accumulate ? 1 : 0);
        slab_finish_kernel<false><<<kFinishBlocks, kSBlock, 0, stream>>>(info, tris, table, P, ws, nullptr);
    }
    return hipGetLastError();
}

}  // namespace voxamd
