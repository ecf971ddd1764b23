// region.h -- write-once table assembly from staged (word, mask) pairs.
//
// The region algorithm splits a voxelization into two launches:
//   1. per workgroup (a chunk of triangles): the voxels or solid run-end
//      markers are merged in an LDS hash keyed by table word, then flushed as
//      (key, mask) pairs into the workgroup's own staging slice -- grouped by
//      output region, with one descriptor (offset, count) per region the
//      flush touches, appended to that region's list;
//   2. per region (2^lg table words): the region's pairs are merged
//      into LDS (OR for surfaces, XOR for solid markers, followed by the row
//      suffix XOR), and the region is written to the table exactly once, so
//      the table needs no zero-fill and no device-scope atomics.
// Pairs that find no room (LDS hash crowded, staging or descriptor list
// full, too many regions in one flush) are applied with a global atomic to
// a scratch table indexed by the same keys, and their region is marked
// dirty; the build pass merges and re-zeros the dirty regions' scratch
// words.  So every pair is applied exactly once whatever the input, and the
// scratch table stays zero between calls.
//
// Key space: key = region << lg | local.
//   RANGE   (Morton surface): key = table word index; a region is an aligned
//           range of 2^lg words, i.e. a brick of the Morton order.
//   ROWS    (linear layout): a region is BY rows in y x BZ rows in z, all x;
//           local = ((z % BZ) * BY + y % BY) * rw + x / 32, rw = G / 32.
//   BRICKS  (Morton solid): as ROWS with BY = 32, BZ = rz; the build pass
//           writes the suffixed rows as Morton bricks.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace voxamd {

enum RegionMode : int { REGION_RANGE = 0, REGION_ROWS = 1, REGION_BRICKS = 2 };

struct RegionPlan {
    int mode;
    uint32_t G;
    int lgG;        // log2 G (G is a power of two)
    int lg;         // log2 key-space words per region
    int lrw;        // log2 words per row (G / 32)
    int lby, lbz;   // log2 rows per region in y and z (ROWS, BRICKS)
    uint32_t nreg;  // regions
    uint32_t dcap;  // descriptors per region
    uint32_t slice; // staged pairs per phase-1 workgroup (its slice starts at blockIdx.x * slice)
    int solid;      // XOR merge + row suffix (else OR)
    int pad;
};

struct RegionWs {
    unsigned* rdc;       // [nreg] descriptors appended per region (zero between calls)
    unsigned* dirty;     // [nreg] 1: the region has scratch words (zero between calls)
    uint2* desc;         // [nreg * dcap] (staging offset, pair count)
    uint2* stage;        // [workgroups * slice] (key, mask)
    uint32_t* scratch;   // [nreg << lg] fallback words (zero between calls)
};

// Capacity overrides (from the VOX_REGION_* flag bits; -1: built-in).
struct RegionKnobs {
    int lg_words = -1, lg_stage = -1, lg_desc = -1;
};

// What a phase-1 kernel needs to stage its pairs (one kernel argument).
struct RegionArgs {
    RegionPlan P;
    RegionWs W;
};

constexpr uint32_t kRegionEmpty = 0xFFFFFFFFu;
constexpr int kRegionProbes = 8;   // LDS hash slots tried before a pair falls back
constexpr int kRegionSlots = 128;  // distinct regions one flush pass groups

// Key of table word (xw = x / 32, y, z) in the ROWS / BRICKS key space.
__device__ __forceinline__ uint32_t region_row_key(const RegionPlan& P, uint32_t xw, uint32_t y, uint32_t z) {
    const uint32_t my = (1u << P.lby) - 1u, mz = (1u << P.lbz) - 1u;
    const uint32_t region = ((z >> P.lbz) << (P.lgG - P.lby)) + (y >> P.lby);
    const uint32_t local = (((((z & mz) << P.lby) | (y & my))) << P.lrw) | xw;
    return (region << P.lg) | local;
}

// A pair that could not be staged: straight into the scratch table.  Out of
// line (the rare path keeps its registers out of the callers' hot loops),
// with scalar arguments: a reference to a kernel argument would make every
// lane copy the structure to its stack.
__device__ __noinline__ void region_fallback_call(uint32_t* scratch, unsigned* dirty, int lg, int solid,
                                                  uint32_t key, uint32_t mask) {
    if (solid)
        atomicXor(scratch + key, mask);
    else
        atomicOr(scratch + key, mask);
    const uint32_t r = key >> lg;
    if (dirty[r] == 0u) dirty[r] = 1u;
}

__device__ __forceinline__ void region_fallback(const RegionPlan& P, const RegionWs& W, uint32_t key, uint32_t mask) {
    region_fallback_call(W.scratch, W.dirty, P.lg, P.solid, key, mask);
}

struct RegionFlushShared {
    uint32_t rkey[kRegionSlots];
    uint32_t rcnt[kRegionSlots];
    uint32_t roff[kRegionSlots];
    uint32_t total;
    int more;  // some pair did not find a region slot in this pass
};

// Flush the pairs a workgroup of NT threads holds, PER per thread in
// registers (key == kRegionEmpty or mask == 0: none).  Pairs are grouped by
// region (up to kRegionSlots regions per pass, further passes for the rest)
// in LDS (lk / lv: NT * PER words each, free for the flush), then copied to
// the workgroup's own staging slice with consecutive stores (no reservation:
// a slice holds NT * PER pairs unless a smaller one is asked for), and
// described to their regions.  Pairs beyond the slice or a full descriptor
// list take the fallback.  Every thread of the workgroup must call it once
// (it synchronises).
template <int NT, int PER>
__device__ __forceinline__ void region_flush(RegionFlushShared& F, const uint32_t (&kk)[PER],
                                             const uint32_t (&vv)[PER], uint32_t* lk, uint32_t* lv,
                                             const RegionPlan& P, const RegionWs& W) {
    bool pending[PER];
#pragma unroll
    for (int u = 0; u < PER; u++) pending[u] = kk[u] != kRegionEmpty && vv[u] != 0u;
    const uint64_t slice0 = (uint64_t)blockIdx.x * P.slice;
    uint32_t used = 0;  // pairs of the slice taken by earlier passes
#pragma unroll 1
    for (;;) {
        for (int i = threadIdx.x; i < kRegionSlots; i += NT) {
            F.rkey[i] = kRegionEmpty;
            F.rcnt[i] = 0u;
        }
        if (threadIdx.x == 0) F.more = 0;
        __syncthreads();
        int slot[PER];
        uint32_t rank[PER];
#pragma unroll
        for (int u = 0; u < PER; u++) {
            slot[u] = -1;
            rank[u] = 0u;
            if (!pending[u]) continue;
            const uint32_t r = kk[u] >> P.lg;
            uint32_t h = (r * 2654435761u) >> 25;  // 7 bits: kRegionSlots
            for (int probe = 0; probe < kRegionSlots; probe++) {
                uint32_t k = F.rkey[h];
                if (k == kRegionEmpty) {
                    k = atomicCAS(&F.rkey[h], kRegionEmpty, r);
                    if (k == kRegionEmpty) k = r;
                }
                if (k == r) {
                    slot[u] = (int)h;
                    rank[u] = atomicAdd(&F.rcnt[h], 1u);
                    break;
                }
                h = (h + 1u) & (kRegionSlots - 1u);
            }
            if (slot[u] < 0) F.more = 1;  // every slot holds another region: next pass
        }
        __syncthreads();
        if (threadIdx.x < 64) {  // exclusive scan of the slot counts (two per lane)
            const int lane = (int)threadIdx.x;
            const uint32_t c0 = F.rcnt[2 * lane], c1 = F.rcnt[2 * lane + 1];
            uint32_t incl = c0 + c1;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t o = __shfl_up(incl, d);
                if (lane >= d) incl += o;
            }
            const uint32_t excl = incl - c0 - c1;
            F.roff[2 * lane] = excl;
            F.roff[2 * lane + 1] = excl + c0;
            const uint32_t total = __shfl(incl, 63);
            if (lane == 0) F.total = total;
        }
        __syncthreads();
        const uint64_t base = slice0 + used;
        const uint32_t total = F.total;  // (F.rcnt is modified below)
        const bool staged = used + total <= P.slice;
        // descriptors first: a region whose list is full takes its pairs as fallbacks
        for (int i = threadIdx.x; i < kRegionSlots; i += NT) {
            const uint32_t c = F.rcnt[i];
            if (!c) continue;
            bool ok = false;
            if (staged) {
                const uint32_t r = F.rkey[i];
                const uint32_t d = atomicAdd(W.rdc + r, 1u);
                if (d < P.dcap) {
                    W.desc[(uint64_t)r * P.dcap + d] = make_uint2((uint32_t)(base + F.roff[i]), c);
                    ok = true;
                }
            }
            if (!ok) F.rcnt[i] = 0u;  // marks the slot's pairs for the fallback
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < PER; u++) {
            if (slot[u] < 0) continue;
            if (F.rcnt[slot[u]]) {
                const uint32_t at = F.roff[slot[u]] + rank[u];
                lk[at] = kk[u];
                lv[at] = vv[u];
            } else {
                region_fallback(P, W, kk[u], vv[u]);
            }
            pending[u] = false;
        }
        const int more = F.more;
        __syncthreads();  // the pass's pairs are in LDS, in staging order
        if (staged) {
            // (the ranges of slots that fell back are copied too, unreferenced)
            for (uint32_t i = threadIdx.x; i < total; i += NT) W.stage[base + i] = make_uint2(lk[i], lv[i]);
            used += total;
        }
        __syncthreads();  // F, lk and lv are rewritten by the next pass
        if (!more) break;
    }
}

}  // namespace voxamd
