// capi.hip -- the C-ABI (vox_capi.h) and the C++ drop-in (voxelize.h).
//
// Host-side rules restated from the reference (Forceflow/cuda_voxelizer v0.6):
//   voxinfo unit        util.h:56-61
//   createMeshBBCube    util.h:80-110
//   table size          main.cpp:190
//   initCuda            util_cuda.cpp:4-42 (device probe; no exit on failure here)
//   voxelize / voxelize_solid launchers voxelize.cu:192-238, voxelize_solid.cu:147-193
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>

#include "../../include/cuda_voxelizer_amd/vox_capi.h"
#include "../../include/cuda_voxelizer_amd/voxelize.h"
#include "vox_kernels.h"
#include "vox_host.h"

namespace voxamd {

static thread_local std::string g_last_error;

vox_status fail(vox_status s, const std::string& msg) {
    g_last_error = msg;
    return s;
}

vox_status hip_fail(hipError_t e, const char* what) {
    return fail(VOX_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Checks shared by every compute entry point.  The grid must be cubic (the
// reference only builds cubic grids, main.cpp:186, and its CPU and GPU paths
// index non-cubic tables differently, cpu_voxelizer.cpp:168 vs voxelize.cu:175);
// the Morton layout needs a power-of-two edge below 2^16 so every code is
// inside the table.
vox_status validate(const vox_info* info, const float* tris, const uint32_t* table, int layout) {
    if (!info) return fail(VOX_ERR_INVALID, "info is NULL");
    const uint32_t G = info->gridsize[0];
    if (G == 0 || info->gridsize[1] != G || info->gridsize[2] != G)
        return fail(VOX_ERR_INVALID, "gridsize must be cubic and non-zero");
    if (G > kMaxGrid) return fail(VOX_ERR_UNSUPPORTED, "grid edge above the supported maximum (8192)");
    for (int k = 0; k < 3; k++) {
        if (!(std::isfinite(info->unit[k]) && info->unit[k] > 0.0f))
            return fail(VOX_ERR_INVALID, "unit must be finite and positive (degenerate bbox?)");
        if (!std::isfinite(info->bbox_min[k])) return fail(VOX_ERR_INVALID, "bbox_min not finite");
    }
    if (layout != VOX_LAYOUT_LINEAR && layout != VOX_LAYOUT_MORTON)
        return fail(VOX_ERR_INVALID, "layout must be VOX_LAYOUT_LINEAR or VOX_LAYOUT_MORTON");
    if (layout == VOX_LAYOUT_MORTON && (G & (G - 1)) != 0)
        return fail(VOX_ERR_UNSUPPORTED, "Morton layout needs a power-of-two grid");
    if (info->n_triangles > 0 && !tris) return fail(VOX_ERR_INVALID, "triangle pointer is NULL");
    if (!table) return fail(VOX_ERR_INVALID, "table pointer is NULL");
    if (info->n_triangles > (uint64_t)kMaxTriangles)
        return fail(VOX_ERR_UNSUPPORTED, "more than 2^31-1 triangles in one call");
    return VOX_OK;
}

// ---- per (device, stream) workspace of the binned pipeline ----------------
// The binned kernels keep the counter/count prefix of the workspace zeroed
// between calls, so a cached workspace is only zero-filled when it is
// (re)allocated.  Calls that share a stream are ordered by it; different
// streams get different workspaces.
struct Workspace {
    void* ptr = nullptr;
    uint64_t bytes = 0;
    uint64_t tiles_cap = 0, n_cap = 0;
};

static std::mutex g_ws_mutex;
static std::map<std::pair<int, hipStream_t>, Workspace> g_ws;

vox_status get_workspace(hipStream_t st, const TilePlan& P, uint64_t n, void** out, WorkspaceLayout* lay) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    Workspace& W = g_ws[std::make_pair(dev, st)];
    if (W.ptr == nullptr || (uint64_t)P.ntiles > W.tiles_cap || n > W.n_cap) {
        // counts are read in uint4 groups up to a multiple of 4096 by the scan
        const uint64_t need = ((uint64_t)P.ntiles + 4095) & ~4095ull;
        const uint64_t tiles_cap = W.tiles_cap > need ? W.tiles_cap : need;
        uint64_t n_cap = W.n_cap > n ? W.n_cap : n;
        if (n_cap < 4096) n_cap = 4096;
        const WorkspaceLayout L = workspace_layout(tiles_cap, n_cap, P.max_tiles_per_tri);
        if (W.ptr) {
            e = hipStreamSynchronize(st);  // the old buffer may still be in use on this stream
            if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
            (void)hipFree(W.ptr);
            W.ptr = nullptr;
        }
        e = hipMalloc(&W.ptr, L.total);
        if (e != hipSuccess) {
            W = Workspace{};
            return fail(VOX_ERR_OOM, std::string("workspace hipMalloc: ") + hipGetErrorString(e));
        }
        e = hipMemset(W.ptr, 0, L.total);
        if (e != hipSuccess) return hip_fail(e, "hipMemset(workspace)");
        W.bytes = L.total;
        W.tiles_cap = tiles_cap;
        W.n_cap = n_cap;
    }
    *out = W.ptr;
    *lay = workspace_layout(W.tiles_cap, W.n_cap, P.max_tiles_per_tri);
    return VOX_OK;
}

// ---- per (device, stream) workspace of the slab pipeline -------------------
// Its misc counters and per-tile counts are left zero by every call, so they
// are zero-filled only when the workspace is (re)allocated.
struct SlabBuffer {
    void* ptr = nullptr;
    SlabLayout L{};
};
static std::mutex g_slab_mutex;
static std::map<std::pair<int, hipStream_t>, SlabBuffer> g_slab;

vox_status get_slab_workspace(hipStream_t st, const SlabPlan& P, uint64_t n, void** out, SlabLayout* lay) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    std::lock_guard<std::mutex> lock(g_slab_mutex);
    SlabBuffer& W = g_slab[std::make_pair(dev, st)];
    if (W.ptr == nullptr || !slab_layout_fits(P, n, W.L)) {
        const uint64_t tiles_cap = W.L.tiles_cap > (uint64_t)P.ntiles ? W.L.tiles_cap : (uint64_t)P.ntiles;
        uint64_t n_cap = W.L.n_cap > n ? W.L.n_cap : n;
        if (n_cap < 4096) n_cap = 4096;
        const SlabLayout L = slab_layout(tiles_cap, n_cap);
        if (!slab_layout_fits(P, n, L)) return fail(VOX_ERR_INVALID, "slab workspace layout does not fit its plan");
        if (W.ptr) {
            e = hipStreamSynchronize(st);  // the old buffer may still be in use on this stream
            if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
            (void)hipFree(W.ptr);
            W.ptr = nullptr;
        }
        e = hipMalloc(&W.ptr, L.total);
        if (e != hipSuccess) {
            W = SlabBuffer{};
            return fail(VOX_ERR_OOM, std::string("slab workspace hipMalloc: ") + hipGetErrorString(e));
        }
        e = hipMemset(W.ptr, 0, L.zero_bytes);
        if (e != hipSuccess) return hip_fail(e, "hipMemset(slab workspace)");
        W.L = L;
    }
    *out = W.ptr;
    *lay = W.L;
    return VOX_OK;
}

// One voxelization.  Algorithms: tile (binned, tile-owned writes; needs a
// plan for this grid/layout), chunk (surface only, LDS word merging + global
// atomics after a fill), marker (solid) and direct (one triangle per lane).
// Without an explicit flag: chunk for surface; tile for a coarse solid mesh,
// direct for a very fine one on a small grid, marker for the others; direct
// when none fits.
vox_status run(const vox_info* info, const float* d_tris, uint32_t* d_table, int layout, unsigned flags,
               uint64_t* d_offrow, void* stream, bool solid) {
    vox_status s = validate(info, d_tris, d_table, layout);
    if (s != VOX_OK) return s;
    const unsigned algo_bits = flags & (VOX_FLAG_ALGO_DIRECT | VOX_FLAG_ALGO_TILE | VOX_FLAG_ALGO_CHUNK |
                                        VOX_FLAG_ALGO_MARKER | VOX_FLAG_ALGO_SLAB | VOX_FLAG_ALGO_REGION);
    if (algo_bits & (algo_bits - 1)) return fail(VOX_ERR_INVALID, "at most one VOX_FLAG_ALGO_* flag");
    if (solid && (flags & VOX_FLAG_ALGO_CHUNK)) return fail(VOX_ERR_INVALID, "the chunk algorithm is surface-only");
    if (!solid && (flags & VOX_FLAG_ALGO_MARKER))
        return fail(VOX_ERR_INVALID, "the marker algorithm is solid-only");
    hipStream_t st = (hipStream_t)stream;
    const uint32_t G = info->gridsize[0];
    const bool accumulate = (flags & VOX_FLAG_ACCUMULATE) != 0;
    TilePlan P;
    const bool have_plan = make_plan((int)G, layout, solid, P, tune_tile_scale(flags)) || make_plan((int)G, layout, solid, P);
    const bool have_marker = solid && marker_supported(G, layout, accumulate);
    SlabPlan SP;
    const int lg_ty = (int)((flags >> 8) & 15u) - 1, lg_tz = (int)((flags >> 12) & 15u) - 1;
    const int split_lg = (int)((flags >> 16) & 31u) - 1;
    const int wg = (int)((flags >> 21) & 3u);
    const bool have_slab = layout == VOX_LAYOUT_LINEAR && (((uintptr_t)d_table) & 15u) == 0 &&
                           make_slab_plan(G, SP, lg_ty, lg_tz, split_lg, wg);
    const bool have_region = region_supported(G, layout, solid);
    enum { DIRECT, TILE, CHUNK, MARKER, SLAB, REGION } algo;
    if (flags & VOX_FLAG_ALGO_DIRECT) algo = DIRECT;
    else if ((flags & VOX_FLAG_ALGO_REGION) && have_region) algo = REGION;
    else if (flags & VOX_FLAG_ALGO_SLAB) algo = have_slab ? SLAB : (have_plan ? TILE : DIRECT);
    else if (flags & VOX_FLAG_ALGO_TILE) algo = have_plan ? TILE : DIRECT;
    else if (flags & VOX_FLAG_ALGO_CHUNK) algo = chunk_supported(G) ? CHUNK : DIRECT;
    else if (flags & VOX_FLAG_ALGO_MARKER) algo = have_marker ? MARKER : (have_plan ? TILE : DIRECT);
    else if (solid && have_plan && coarse_mesh(info->n_triangles, G)) algo = TILE;
    else if (solid && prefer_direct_solid(info->n_triangles, G)) algo = DIRECT;
    else if (solid) algo = have_marker ? MARKER : (have_plan ? TILE : DIRECT);
    else algo = chunk_supported(G) ? CHUNK : (have_plan ? TILE : DIRECT);
    hipError_t e;
    if (algo == REGION) {
        const RegionKnobs K = region_knobs(flags);
        if (solid)
            e = launch_solid_region(*info, d_tris, d_table, layout, accumulate, (unsigned long long*)d_offrow, st,
                                    tune_big(flags), K);
        else
            e = launch_surface_region(*info, d_tris, d_table, layout, accumulate, st, tune_big(flags),
                                      tune_waves(flags), K);
        if (e != hipSuccess) return hip_fail(e, "region kernels launch");
        return VOX_OK;
    }
    if (algo == SLAB) {
        void* ws = nullptr;
        SlabLayout L;
        s = get_slab_workspace(st, SP, info->n_triangles, &ws, &L);
        if (s != VOX_OK) return s;
        e = launch_slab(*info, d_tris, d_table, SP, solid, accumulate, (unsigned long long*)d_offrow, ws, L, st);
        if (e != hipSuccess) return hip_fail(e, "slab kernels launch");
        return VOX_OK;
    }
    if (algo == MARKER) {
        e = launch_solid_marker(*info, d_tris, d_table, layout, accumulate, (unsigned long long*)d_offrow, st,
                                tune_big(flags));
        if (e != hipSuccess) return hip_fail(e, "marker kernels launch");
        return VOX_OK;
    }
    if (algo == TILE) {
        void* ws = nullptr;
        WorkspaceLayout L;
        s = get_workspace(st, P, info->n_triangles, &ws, &L);
        if (s != VOX_OK) return s;
        e = launch_tiled(*info, d_tris, d_table, P, solid, layout, accumulate, ws, L,
                         (unsigned long long*)d_offrow, st);
        if (e != hipSuccess) return hip_fail(e, "tile kernels launch");
        return VOX_OK;
    }
    if (!accumulate) {
        e = launch_fill_zero(*info, d_table, st);
        if (e != hipSuccess) return hip_fail(e, "table zero-fill");
    }
    if (algo == CHUNK)
        e = launch_surface_chunk(*info, d_tris, d_table, layout, st, tune_big(flags), tune_waves(flags));
    else if (solid)
        e = launch_solid_direct(*info, d_tris, d_table, layout, (unsigned long long*)d_offrow, st);
    else
        e = launch_surface_direct(*info, d_tris, d_table, layout, st);
    if (e != hipSuccess) return hip_fail(e, "kernel launch");
    return VOX_OK;
}

}  // namespace voxamd

using namespace voxamd;

extern "C" {

const char* vox_last_error(void) { return g_last_error.c_str(); }

const char* vox_version(void) { return "cuda_voxelizer_amd 0.1 (gfx950)"; }

vox_status vox_init(int device) {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count < 1) return fail(VOX_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= count) return fail(VOX_ERR_INVALID, "device index out of range");
    e = hipSetDevice(device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    // the device address of vox_solid_noalloc's static scratch on this GPU
    e = static_scratch_init(device, nullptr);
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    return VOX_OK;
}

size_t vox_table_bytes(uint32_t gx, uint32_t gy, uint32_t gz) {
    const uint64_t voxels = (uint64_t)gx * gy * gz;
    return (size_t)((voxels + 31) / 32) * 4;
}

vox_status vox_make_info(const float bbox_min[3], const float bbox_max[3], uint32_t grid, uint64_t n_triangles,
                         vox_info* out) {
    if (!bbox_min || !bbox_max || !out) return fail(VOX_ERR_INVALID, "NULL argument");
    if (grid == 0) return fail(VOX_ERR_INVALID, "grid must be positive");
    std::memset(out, 0, sizeof *out);
    for (int k = 0; k < 3; k++) {
        out->bbox_min[k] = bbox_min[k];
        out->bbox_max[k] = bbox_max[k];
        out->gridsize[k] = grid;
        // unit.k = (bbox.max.k - bbox.min.k) / float(gridsize.k)
        out->unit[k] = (bbox_max[k] - bbox_min[k]) / (float)grid;
    }
    out->n_triangles = n_triangles;
    return VOX_OK;
}

vox_status vox_bbox_cube(const float in_min[3], const float in_max[3], float out_min[3], float out_max[3]) {
    if (!in_min || !in_max || !out_min || !out_max) return fail(VOX_ERR_INVALID, "NULL argument");
    float len[3], mn[3], mx[3];
    for (int k = 0; k < 3; k++) {
        len[k] = in_max[k] - in_min[k];
        mn[k] = in_min[k];
        mx[k] = in_max[k];
    }
    // std::max(a, b) == (a < b) ? b : a
    const float m_yz = (len[1] < len[2]) ? len[2] : len[1];
    const float max_length = (len[0] < m_yz) ? m_yz : len[0];
    for (int k = 0; k < 3; k++) {
        if (max_length != len[k]) {
            const float half = (max_length - len[k]) / 2.0f;
            mn[k] = in_min[k] - half;
            mx[k] = in_max[k] + half;
        }
    }
    for (int k = 0; k < 3; k++) {
        const float eps = (mx[k] - mn[k]) / 10001.0f;
        out_min[k] = mn[k] - eps;
        out_max[k] = mx[k] + eps;
    }
    return VOX_OK;
}

vox_status vox_query_tile_plan(uint32_t grid, int layout, int solid, int32_t out[8]) {
    if (!out) return fail(VOX_ERR_INVALID, "NULL argument");
    TilePlan P;
    if (grid == 0 || grid > kMaxGrid || !make_plan((int)grid, layout, solid != 0, P))
        return fail(VOX_ERR_UNSUPPORTED, "no tile plan for this grid/layout/mode");
    out[0] = P.kind; out[1] = P.tx; out[2] = P.ty; out[3] = P.tz;
    out[4] = P.ntiles; out[5] = P.tile_words; out[6] = P.row_words; out[7] = P.max_tiles_per_tri;
    return VOX_OK;
}

vox_status vox_surface(const vox_info* info, const float* d_tris, uint32_t* d_table, int layout, unsigned flags,
                       void* stream) {
    return run(info, d_tris, d_table, layout, flags, nullptr, stream, false);
}

vox_status vox_solid(const vox_info* info, const float* d_tris, uint32_t* d_table, int layout, unsigned flags,
                     uint64_t* d_offrow, void* stream) {
    return run(info, d_tris, d_table, layout, flags, d_offrow, stream, true);
}

vox_status vox_solid_noalloc(const vox_info* info, const float* d_tris, uint32_t* d_table, uint64_t* d_offrow,
                             void* stream) {
    vox_status s = validate(info, d_tris, d_table, VOX_LAYOUT_LINEAR);
    if (s != VOX_OK) return s;
    const uint32_t G = info->gridsize[0];
    if (!marker_supported(G, VOX_LAYOUT_LINEAR, false))
        return fail(VOX_ERR_UNSUPPORTED,
                    "vox_solid_noalloc: the grid edge must be a multiple of 32 between 32 and 16384 "
                    "(its fixed scratch holds only the spill counters and records)");
    const hipError_t e = launch_solid_marker_static(*info, d_tris, d_table, (unsigned long long*)d_offrow,
                                                    (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "vox_solid_noalloc launch");
    return VOX_OK;
}

size_t vox_noalloc_scratch_bytes(void) { return static_scratch_bytes(); }

}  // extern "C"

// ---------------------------------------------------------------------------
// C++ drop-in (voxelize.h)
// ---------------------------------------------------------------------------
static bool g_verbose = true;

void voxelize_set_verbose(bool verbose) { g_verbose = verbose; }

AABox<float3> createMeshBBCube(AABox<float3> box) {
    const float in_min[3] = {box.min.x, box.min.y, box.min.z};
    const float in_max[3] = {box.max.x, box.max.y, box.max.z};
    float out_min[3], out_max[3];
    vox_bbox_cube(in_min, in_max, out_min, out_max);
    return AABox<float3>(make_float3(out_min[0], out_min[1], out_min[2]),
                         make_float3(out_max[0], out_max[1], out_max[2]));
}

namespace {

// checkCudaErrors-style: print and exit(EXIT_FAILURE) (helper_cuda.h:566-579).
void die_on(vox_status s, const char* what) {
    if (s == VOX_OK) return;
    std::fprintf(stderr, "[Err] %s failed: %s\n", what, vox_last_error());
    std::exit(EXIT_FAILURE);
}

void die_on_hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return;
    std::fprintf(stderr, "[Err] %s failed: %s\n", what, hipGetErrorString(e));
    std::exit(EXIT_FAILURE);
}

template <typename F>
void run_timed(F&& f) {
    hipEvent_t start, stop;
    die_on_hip(hipEventCreate(&start), "hipEventCreate");
    die_on_hip(hipEventCreate(&stop), "hipEventCreate");
    die_on_hip(hipEventRecord(start, 0), "hipEventRecord");
    f();
    die_on_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
    die_on_hip(hipEventRecord(stop, 0), "hipEventRecord");
    die_on_hip(hipEventSynchronize(stop), "hipEventSynchronize");
    float ms = 0.0f;
    die_on_hip(hipEventElapsedTime(&ms, start, stop), "hipEventElapsedTime");
    if (g_verbose) std::printf("[Perf] Voxelization GPU time: %.1f ms\n", ms);
    die_on_hip(hipEventDestroy(start), "hipEventDestroy");
    die_on_hip(hipEventDestroy(stop), "hipEventDestroy");
}

const vox_info* as_capi(const voxinfo& v) { return reinterpret_cast<const vox_info*>(&v); }

}  // namespace

void voxelize(const voxinfo& v, float* triangle_data, unsigned int* vtable, bool morton_code) {
    run_timed([&] {
        die_on(vox_surface(as_capi(v), triangle_data, vtable,
                           morton_code ? VOX_LAYOUT_MORTON : VOX_LAYOUT_LINEAR, VOX_FLAG_ACCUMULATE, nullptr),
               "voxelize");
    });
}

void voxelize(const voxinfo& v, float* triangle_data, unsigned int* vtable, bool /*useThrustPath*/,
              bool morton_code) {
    voxelize(v, triangle_data, vtable, morton_code);
}

void voxelize_solid(const voxinfo& v, float* triangle_data, unsigned int* vtable, bool morton_code) {
    run_timed([&] {
        die_on(vox_solid(as_capi(v), triangle_data, vtable, morton_code ? VOX_LAYOUT_MORTON : VOX_LAYOUT_LINEAR,
                         VOX_FLAG_ACCUMULATE, nullptr, nullptr),
               "voxelize_solid");
    });
}
