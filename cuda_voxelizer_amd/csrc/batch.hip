// batch.hip -- many voxelizations in one graph launch (vox_batch_* C-ABI).
//
// vox_batch_create sorts the jobs into algorithm classes (fill, chunk,
// marker, direct surface/solid, tile surface/solid), uploads one device array of job
// descriptors per class together with the per-job prefix of each kernel's
// work units, allocates one workspace, and captures the whole pipeline --
// every kernel processing all jobs of its class at once -- into a HIP graph.
// vox_batch_launch then costs one hipGraphLaunch per batch, however many
// meshes it holds.
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/cuda_voxelizer_amd/vox_capi.h"
#include "vox_host.h"
#include "vox_kernels.h"

namespace voxamd {

struct JobClass {
    std::vector<ItemDev> items;
    std::vector<uint32_t> base_a, base_b, base_c;  // prefixes (class dependent)
    ItemMap m_a{}, m_b{}, m_c{};
    uint32_t total_a = 0, total_b = 0, total_c = 0;
    uint32_t max_tile_words = 0;
    bool loop = false;  // chunk class: some job takes several chunks per workgroup
    size_t off_items = 0, off_a = 0, off_b = 0, off_c = 0;
};

}  // namespace voxamd

struct vox_batch {
    int device = 0;
    bool use_graph = true;
    hipStream_t cap_stream = nullptr;
    hipStream_t side[4] = {};  // capture only: tile chain, marker chain, direct classes, generic chunk classes
    hipEvent_t ev[6] = {};     // capture only: fork, tables zeroed, and the four side branches' ends
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void* dev_mem = nullptr;
    void* ws = nullptr;
    void* mws = nullptr;  // marker class: spill records and counters
    voxamd::WorkspaceLayout L{};
    // chunk / chunkw (6 waves per SIMD): one class per layout specialisation (ChunkSpec)
    voxamd::JobClass fill, chunk[3], chunkw[3], marker, dsurf, dsolid, tsurf, tsolid;
    int count = 0;
    uint64_t total_triangles = 0;
};

namespace voxamd {

static void add_prefix(std::vector<uint32_t>& base, uint32_t& total, uint64_t units) {
    if (base.empty()) base.push_back(0);
    total += (uint32_t)units;
    base.push_back(total);
}

// The classes of a batch, by the launch that runs them.  Every job is in one
// algorithm class; the fill class zeroes the tables of the chunk, marker and
// direct jobs that do not accumulate (tile jobs write every word themselves).
static hipError_t launch_fill_class(vox_batch* b, hipStream_t s) {
    const ItemDev dummy{};
    if (!b->fill.total_a) return hipSuccess;
    return launch_fill_jobs(dummy, b->fill.m_a, b->fill.total_a, s);
}

// Which chunk classes a launch covers: all, the layout-generic ones (the
// chunk loop of small dense tables among them) or the layout-specialised ones.
enum ChunkClassSet { kAllChunkClasses, kGenericChunkClasses, kSpecialisedChunkClasses };

static hipError_t launch_chunk_classes(vox_batch* b, hipStream_t s, ChunkClassSet which = kAllChunkClasses) {
    const ItemDev dummy{};
    hipError_t e = hipSuccess;
    for (int sp = 0; sp < 3; sp++) {
        if ((which == kGenericChunkClasses && sp != kSpecAny) || (which == kSpecialisedChunkClasses && sp == kSpecAny))
            continue;
        const JobClass& c5 = b->chunk[sp];
        const JobClass& c6 = b->chunkw[sp];
        if (c5.total_a && (e = launch_chunk_jobs(dummy, c5.m_a, c5.total_a, c5.loop, s, 5, sp)) != hipSuccess)
            return e;
        if (c6.total_a && (e = launch_chunk_jobs(dummy, c6.m_a, c6.total_a, c6.loop, s, 6, sp)) != hipSuccess)
            return e;
    }
    return hipSuccess;
}

static hipError_t launch_marker_class(vox_batch* b, hipStream_t s) {
    if (b->marker.items.empty()) return hipSuccess;
    return launch_marker_jobs(b->marker.m_a, b->marker.m_b, b->marker.m_c, b->marker.total_a, b->marker.total_b,
                              b->mws, s);
}

static hipError_t launch_direct_classes(vox_batch* b, hipStream_t s) {
    const ItemDev dummy{};
    hipError_t e = hipSuccess;
    if (b->dsurf.total_a && (e = launch_direct_jobs(dummy, b->dsurf.m_a, b->dsurf.total_a, false, s)) != hipSuccess)
        return e;
    if (b->dsolid.total_a && (e = launch_direct_jobs(dummy, b->dsolid.m_a, b->dsolid.total_a, true, s)) != hipSuccess)
        return e;
    return hipSuccess;
}

// The two tile classes share the binning workspace: one after the other.
static hipError_t launch_tile_classes(vox_batch* b, hipStream_t s) {
    const ItemDev dummy{};
    hipError_t e = hipSuccess;
    if (!b->tsurf.items.empty() &&
        (e = launch_tiled_jobs(dummy, b->tsurf.m_a, b->tsurf.m_b, b->tsurf.m_c, b->tsurf.total_a, b->tsurf.total_b,
                               false, b->tsurf.max_tile_words, b->ws, b->L, s)) != hipSuccess)
        return e;
    if (!b->tsolid.items.empty() &&
        (e = launch_tiled_jobs(dummy, b->tsolid.m_a, b->tsolid.m_b, b->tsolid.m_c, b->tsolid.total_a,
                               b->tsolid.total_b, true, b->tsolid.max_tile_words, b->ws, b->L, s)) != hipSuccess)
        return e;
    return hipSuccess;
}

// Every class in stream order (eager launches and runners).
static hipError_t launch_all(vox_batch* b, hipStream_t s) {
    hipError_t e = hipSuccess;
    if ((e = launch_fill_class(b, s)) != hipSuccess) return e;
    if ((e = launch_chunk_classes(b, s)) != hipSuccess) return e;
    if ((e = launch_marker_class(b, s)) != hipSuccess) return e;
    if ((e = launch_direct_classes(b, s)) != hipSuccess) return e;
    return launch_tile_classes(b, s);
}

// For graph capture on `m`: classes that touch different jobs run as
// concurrent branches, so one class's tail (its last workgroups) and the
// small latency-bound kernels (the tile chain's scans, the chunk loop of
// small dense tables, the marker chain's spill pass) overlap the others
// instead of leaving the GPU idle between kernels.  Nodes are created in the
// order the critical path wants them dispatched: the fill, then the
// layout-specialised chunk classes (the bulk of most batches) on `m`, then
// the branches -- the tile chain (no fill; it forks from before the fill),
// the marker chain, the direct classes and the layout-generic chunk classes
// (after the fill).  `m` joins every branch at the end.  Tables of different
// jobs do not overlap, so the branches share no data (the out-of-row
// counters they add to are atomic).
static hipError_t launch_forked(vox_batch* b, hipStream_t m) {
    hipError_t e = hipSuccess;
#define VOX_TRY(x) \
    if ((e = (x)) != hipSuccess) return e
    VOX_TRY(hipEventRecord(b->ev[0], m));
    VOX_TRY(launch_fill_class(b, m));
    VOX_TRY(hipEventRecord(b->ev[1], m));
    VOX_TRY(launch_chunk_classes(b, m, kSpecialisedChunkClasses));
    VOX_TRY(hipStreamWaitEvent(b->side[0], b->ev[0], 0));
    VOX_TRY(launch_tile_classes(b, b->side[0]));
    VOX_TRY(hipEventRecord(b->ev[2], b->side[0]));
    VOX_TRY(hipStreamWaitEvent(b->side[1], b->ev[1], 0));
    VOX_TRY(launch_marker_class(b, b->side[1]));
    VOX_TRY(hipEventRecord(b->ev[3], b->side[1]));
    VOX_TRY(hipStreamWaitEvent(b->side[2], b->ev[1], 0));
    VOX_TRY(launch_direct_classes(b, b->side[2]));
    VOX_TRY(hipEventRecord(b->ev[4], b->side[2]));
    VOX_TRY(hipStreamWaitEvent(b->side[3], b->ev[1], 0));
    VOX_TRY(launch_chunk_classes(b, b->side[3], kGenericChunkClasses));
    VOX_TRY(hipEventRecord(b->ev[5], b->side[3]));
    for (int k = 2; k < 6; k++) VOX_TRY(hipStreamWaitEvent(m, b->ev[k], 0));
#undef VOX_TRY
    return hipSuccess;
}

static void free_batch(vox_batch* b) {
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
    for (hipStream_t& st : b->side)
        if (st) (void)hipStreamDestroy(st);
    for (hipEvent_t& ev : b->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (b->dev_mem) (void)hipFree(b->dev_mem);
    if (b->ws) (void)hipFree(b->ws);
    if (b->mws) (void)hipFree(b->mws);
    delete b;
}

// Host side of a batch: jobs sorted into algorithm classes, the packed
// device image of the descriptor and prefix arrays, and the workspace needed.
struct BatchPlan {
    std::vector<char> host;
    bool ws_needed = false;
    WorkspaceLayout L{};
};

static void reset_classes(vox_batch* b) {
    JobClass* classes[] = {&b->fill,      &b->chunk[0],  &b->chunk[1], &b->chunk[2], &b->chunkw[0], &b->chunkw[1],
                           &b->chunkw[2], &b->marker,   &b->dsurf,    &b->dsolid,   &b->tsurf,     &b->tsolid};
    for (JobClass* c : classes) *c = JobClass{};
    b->count = 0;
    b->total_triangles = 0;
}

static vox_status plan_batch(const vox_batch_item* items, int count, vox_batch* b, BatchPlan& plan) {
    if (count < 0 || (count > 0 && !items)) return fail(VOX_ERR_INVALID, "bad item array");
    reset_classes(b);
    b->count = count;
    uint64_t tri_sum_surf = 0, tri_sum_solid = 0;
    for (int i = 0; i < count; i++) {
        const vox_batch_item& it = items[i];
        vox_status s = validate(&it.info, it.d_tris, it.d_table, it.layout);
        if (s != VOX_OK) return fail(s, "item " + std::to_string(i) + ": " + vox_last_error());
        const unsigned algo =
            it.flags & (VOX_FLAG_ALGO_DIRECT | VOX_FLAG_ALGO_TILE | VOX_FLAG_ALGO_CHUNK | VOX_FLAG_ALGO_MARKER);
        const bool solid = it.solid != 0;
        if ((algo & (algo - 1)) || (solid && (algo & VOX_FLAG_ALGO_CHUNK)) || (!solid && (algo & VOX_FLAG_ALGO_MARKER)))
            return fail(VOX_ERR_INVALID, "item " + std::to_string(i) + ": bad algorithm flags");
        const uint32_t G = it.info.gridsize[0];
        const bool accumulate = (it.flags & VOX_FLAG_ACCUMULATE) != 0;
        TilePlan P;
        const bool have_plan =
            make_plan((int)G, it.layout, solid, P, tune_tile_scale(it.flags)) || make_plan((int)G, it.layout, solid, P);
        // the marker class tags spill records with the job index (below 2^22)
        const bool have_marker = solid && marker_batch_supported(G, it.layout, accumulate) && i < (1 << 22);
        enum { DIRECT, TILE, CHUNK, MARKER } a;
        if (algo & VOX_FLAG_ALGO_DIRECT) a = DIRECT;
        else if (algo & VOX_FLAG_ALGO_TILE) a = have_plan ? TILE : DIRECT;
        else if (algo & VOX_FLAG_ALGO_CHUNK) a = chunk_supported(G) ? CHUNK : DIRECT;
        else if (algo & VOX_FLAG_ALGO_MARKER) a = have_marker ? MARKER : (have_plan ? TILE : DIRECT);
        else if (solid && prefer_direct_solid(it.info.n_triangles, G)) a = DIRECT;
        else if (solid && have_marker && !(have_plan && coarse_mesh(it.info.n_triangles, G))) a = MARKER;
        else if (solid) a = have_plan ? TILE : DIRECT;
        else a = chunk_supported(G) ? CHUNK : (have_plan ? TILE : DIRECT);
        ItemDev d{};
        d.info = it.info;
        d.tris = it.d_tris;
        d.table = it.d_table;
        d.offrow = (unsigned long long*)it.d_offrow;
        d.layout = it.layout;
        d.solid = solid ? 1 : 0;
        d.accumulate = accumulate ? 1 : 0;
        d.big = tune_big(it.flags);
        const uint64_t n = it.info.n_triangles;
        b->total_triangles += n;
        if (a != TILE && !accumulate) {
            b->fill.items.push_back(d);
            add_prefix(b->fill.base_a, b->fill.total_a, fill_units(G));
        }
        if (a == MARKER) {
            if ((uint64_t)b->marker.total_c + n > 0xFFFFFFFFull)
                return fail(VOX_ERR_UNSUPPORTED, "more than 2^32-1 triangles in one marker batch class");
            b->marker.items.push_back(d);
            add_prefix(b->marker.base_a, b->marker.total_a, marker_block_units(n));
            add_prefix(b->marker.base_b, b->marker.total_b, marker_row_units(G));
            add_prefix(b->marker.base_c, b->marker.total_c, n);  // triangles (spill overflow pass)
        } else if (a == CHUNK) {
            d.span = chunk_span(G, n);
            d.parts = chunk_parts(G, n);
            // jobs that loop over chunks run in the 4-wave form (any layout): they stay in the 5-wave class;
            // a layout-specialised kernel runs at six waves per SIMD under either bound, so only
            // layout-generic jobs are split by the waves knob (each class is one more kernel)
            const bool loop = d.span > 1u && d.parts <= 1u;
            const int sp = loop ? (int)kSpecAny : chunk_spec(it.layout, G);
            const bool wide = tune_waves(it.flags) == 6 && !loop && sp == kSpecAny;
            JobClass& c = wide ? b->chunkw[sp] : b->chunk[sp];
            if (loop) c.loop = true;
            c.items.push_back(d);
            add_prefix(c.base_a, c.total_a, chunk_units(n, d.span, d.parts));
        } else if (a == DIRECT) {
            JobClass& c = solid ? b->dsolid : b->dsurf;
            c.items.push_back(d);
            add_prefix(c.base_a, c.total_a, direct_units(n));
        } else {
            JobClass& c = solid ? b->tsolid : b->tsurf;
            uint64_t& tri_sum = solid ? tri_sum_solid : tri_sum_surf;
            d.P = P;
            d.tile_base = c.total_b;
            d.tri_base = tri_sum;
            if (tri_sum + n > 0xFFFFFFFFull)
                return fail(VOX_ERR_UNSUPPORTED, "more than 2^32-1 triangles in one tiled batch class");
            tri_sum += n;
            c.items.push_back(d);
            add_prefix(c.base_a, c.total_a, (n + 255) / 256);  // bin blocks
            add_prefix(c.base_b, c.total_b, (uint64_t)P.ntiles);
            add_prefix(c.base_c, c.total_c, n);                 // triangles
            if ((uint32_t)P.tile_words > c.max_tile_words) c.max_tile_words = (uint32_t)P.tile_words;
        }
    }
    // one device allocation: per class the job array and its prefix arrays
    JobClass* classes[] = {&b->fill,      &b->chunk[0],  &b->chunk[1], &b->chunk[2], &b->chunkw[0], &b->chunkw[1],
                           &b->chunkw[2], &b->marker,   &b->dsurf,    &b->dsolid,   &b->tsurf,     &b->tsolid};
    size_t total = 0;
    auto reserve = [&](size_t bytes) {
        const size_t off = total;
        total = (total + bytes + 255) & ~(size_t)255;
        return off;
    };
    for (JobClass* c : classes) {
        c->off_items = reserve(c->items.size() * sizeof(ItemDev));
        c->off_a = reserve(c->base_a.size() * 4);
        c->off_b = reserve(c->base_b.size() * 4);
        c->off_c = reserve(c->base_c.size() * 4);
    }
    std::vector<char>& host = plan.host;
    host.assign(total > 0 ? total : 1, 0);
    for (JobClass* c : classes) {
        if (!c->items.empty()) std::memcpy(host.data() + c->off_items, c->items.data(), c->items.size() * sizeof(ItemDev));
        if (!c->base_a.empty()) std::memcpy(host.data() + c->off_a, c->base_a.data(), c->base_a.size() * 4);
        if (!c->base_b.empty()) std::memcpy(host.data() + c->off_b, c->base_b.data(), c->base_b.size() * 4);
        if (!c->base_c.empty()) std::memcpy(host.data() + c->off_c, c->base_c.data(), c->base_c.size() * 4);
    }
    if (!b->tsurf.items.empty() || !b->tsolid.items.empty()) {
        const uint64_t tiles = b->tsurf.total_b > b->tsolid.total_b ? b->tsurf.total_b : b->tsolid.total_b;
        const uint64_t tris = tri_sum_surf > tri_sum_solid ? tri_sum_surf : tri_sum_solid;
        plan.ws_needed = true;
        plan.L = workspace_layout((tiles + 4095) & ~4095ull, tris > 4096 ? tris : 4096, 8);
    }
    return VOX_OK;
}

// Point every class's job maps at a device copy of the packed image.
static void bind_maps(vox_batch* b, void* dev) {
    char* dm = (char*)dev;
    JobClass* classes[] = {&b->fill,      &b->chunk[0],  &b->chunk[1], &b->chunk[2], &b->chunkw[0], &b->chunkw[1],
                           &b->chunkw[2], &b->marker,   &b->dsurf,    &b->dsolid,   &b->tsurf,     &b->tsolid};
    for (JobClass* c : classes) {
        const ItemDev* it = (const ItemDev*)(dm + c->off_items);
        const int ni = (int)c->items.size();
        c->m_a = ItemMap{it, (const uint32_t*)(dm + c->off_a), ni, 0};
        c->m_b = ItemMap{it, (const uint32_t*)(dm + c->off_b), ni, 0};
        c->m_c = ItemMap{it, (const uint32_t*)(dm + c->off_c), ni, 0};
    }
}

}  // namespace voxamd

using namespace voxamd;

extern "C" {

vox_status vox_batch_create(const vox_batch_item* items, int count, unsigned flags, vox_batch** out) {
    if (!out) return fail(VOX_ERR_INVALID, "out is NULL");
    *out = nullptr;
    vox_batch* b = new (std::nothrow) vox_batch();
    if (!b) return fail(VOX_ERR_OOM, "host allocation");
    b->use_graph = !(flags & VOX_BATCH_NO_GRAPH);
    hipError_t e = hipGetDevice(&b->device);
    if (e != hipSuccess) {
        free_batch(b);
        return hip_fail(e, "hipGetDevice");
    }
    BatchPlan plan;
    vox_status s = plan_batch(items, count, b, plan);
    if (s != VOX_OK) {
        free_batch(b);
        return s;
    }
    e = hipMalloc(&b->dev_mem, plan.host.size());
    if (e != hipSuccess) {
        free_batch(b);
        return fail(VOX_ERR_OOM, std::string("batch descriptors: ") + hipGetErrorString(e));
    }
    e = hipMemcpy(b->dev_mem, plan.host.data(), plan.host.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        free_batch(b);
        return hip_fail(e, "batch descriptor upload");
    }
    bind_maps(b, b->dev_mem);
    if (!b->marker.items.empty()) {
        e = hipMalloc(&b->mws, marker_ws_bytes());
        if (e == hipSuccess) e = hipMemset(b->mws, 0, marker_ws_bytes());  // counters start at zero
        if (e != hipSuccess) {
            free_batch(b);
            return fail(VOX_ERR_OOM, std::string("batch marker workspace: ") + hipGetErrorString(e));
        }
    }
    if (plan.ws_needed) {
        b->L = plan.L;
        e = hipMalloc(&b->ws, b->L.total);
        if (e == hipSuccess) e = hipMemset(b->ws, 0, b->L.total);
        if (e != hipSuccess) {
            free_batch(b);
            return fail(VOX_ERR_OOM, std::string("batch workspace: ") + hipGetErrorString(e));
        }
    }
    if (b->use_graph) {
        e = hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking);
        for (hipStream_t& st : b->side)
            if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        for (hipEvent_t& ev : b->ev)
            if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e == hipSuccess) e = hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            hipError_t le = launch_forked(b, b->cap_stream);
            e = hipStreamEndCapture(b->cap_stream, &b->graph);
            if (le != hipSuccess) e = le;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            free_batch(b);
            return hip_fail(e, "batch graph capture");
        }
    }
    *out = b;
    return VOX_OK;
}

vox_status vox_batch_launch(vox_batch* b, void* stream) {
    if (!b) return fail(VOX_ERR_INVALID, "batch is NULL");
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = b->use_graph ? hipGraphLaunch(b->exec, st) : launch_all(b, st);
    if (e != hipSuccess) return hip_fail(e, "batch launch");
    return VOX_OK;
}

vox_status vox_batch_info(const vox_batch* b, int64_t out[8]) {
    if (!b || !out) return fail(VOX_ERR_INVALID, "NULL argument");
    out[0] = b->count;
    out[1] = (int64_t)b->total_triangles;
    out[2] = 0;
    for (int sp = 0; sp < 3; sp++) out[2] += (int64_t)(b->chunk[sp].items.size() + b->chunkw[sp].items.size());
    out[3] = (int64_t)b->tsurf.items.size();
    out[4] = (int64_t)b->tsolid.items.size();
    out[5] = (int64_t)(b->dsurf.items.size() + b->dsolid.items.size());
    out[6] = b->use_graph ? 1 : 0;
    out[7] = (int64_t)b->L.total;
    return VOX_OK;
}

vox_status vox_batch_destroy(vox_batch* b) {
    if (!b) return VOX_OK;
    (void)hipDeviceSynchronize();  // launches of this batch may still be in flight on caller streams
    free_batch(b);
    return VOX_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// vox_runner: a re-fillable eager batch bound to one stream (serving)
// ---------------------------------------------------------------------------
struct vox_runner {
    vox_batch b;                 // classes, job maps, workspace (eager launches)
    hipStream_t stream = nullptr;
    void* dev_desc = nullptr;
    size_t dev_cap = 0;
    void* pinned[2] = {nullptr, nullptr};
    size_t pinned_cap[2] = {0, 0};
    hipEvent_t ev[2] = {nullptr, nullptr};
    int slot = 0;
    uint64_t tiles_cap = 0, n_cap = 0;
    uint64_t launches = 0;
};

extern "C" {

vox_status vox_runner_create(void* stream, vox_runner** out) {
    if (!out) return fail(VOX_ERR_INVALID, "out is NULL");
    *out = nullptr;
    vox_runner* r = new (std::nothrow) vox_runner();
    if (!r) return fail(VOX_ERR_OOM, "host allocation");
    r->stream = (hipStream_t)stream;
    r->b.use_graph = false;
    hipError_t e = hipGetDevice(&r->b.device);
    for (int i = 0; i < 2 && e == hipSuccess; i++) e = hipEventCreateWithFlags(&r->ev[i], hipEventDisableTiming);
    if (e != hipSuccess) {
        vox_runner_destroy(r);
        return hip_fail(e, "runner setup");
    }
    *out = r;
    return VOX_OK;
}

vox_status vox_runner_launch(vox_runner* r, const vox_batch_item* items, int count) {
    if (!r) return fail(VOX_ERR_INVALID, "runner is NULL");
    BatchPlan plan;
    vox_status s = plan_batch(items, count, &r->b, plan);
    if (s != VOX_OK) return s;
    hipError_t e;
    // grow-only device buffers; growing waits for this runner's stream
    const uint64_t tiles = r->b.tsurf.total_b > r->b.tsolid.total_b ? r->b.tsurf.total_b : r->b.tsolid.total_b;
    uint64_t tris = 0;
    if (!r->b.tsurf.base_c.empty()) tris = r->b.tsurf.base_c.back();
    if (!r->b.tsolid.base_c.empty() && r->b.tsolid.base_c.back() > tris) tris = r->b.tsolid.base_c.back();
    const bool grow_ws = plan.ws_needed && (tiles > r->tiles_cap || tris > r->n_cap || !r->b.ws);
    const bool grow_desc = plan.host.size() > r->dev_cap;
    if (grow_ws || grow_desc) {
        e = hipStreamSynchronize(r->stream);
        if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    }
    if (grow_ws) {
        if (r->b.ws) (void)hipFree(r->b.ws);
        r->b.ws = nullptr;
        r->tiles_cap = ((tiles > r->tiles_cap ? tiles : r->tiles_cap) + 4095) & ~4095ull;
        r->n_cap = tris > r->n_cap ? tris : r->n_cap;
        if (r->n_cap < 4096) r->n_cap = 4096;
        r->b.L = workspace_layout(r->tiles_cap, r->n_cap, 8);
        e = hipMalloc(&r->b.ws, r->b.L.total);
        if (e == hipSuccess) e = hipMemset(r->b.ws, 0, r->b.L.total);
        if (e != hipSuccess) {
            r->tiles_cap = r->n_cap = 0;
            return fail(VOX_ERR_OOM, std::string("runner workspace: ") + hipGetErrorString(e));
        }
    }
    if (!r->b.marker.items.empty() && !r->b.mws) {  // fixed size, allocated once
        e = hipMalloc(&r->b.mws, marker_ws_bytes());
        if (e == hipSuccess) e = hipMemset(r->b.mws, 0, marker_ws_bytes());  // counters start at zero
        if (e != hipSuccess) {
            if (r->b.mws) (void)hipFree(r->b.mws);
            r->b.mws = nullptr;
            return fail(VOX_ERR_OOM, std::string("runner marker workspace: ") + hipGetErrorString(e));
        }
    }
    if (grow_desc) {
        if (r->dev_desc) (void)hipFree(r->dev_desc);
        r->dev_cap = plan.host.size() * 2;
        e = hipMalloc(&r->dev_desc, r->dev_cap);
        if (e != hipSuccess) {
            r->dev_cap = 0;
            r->dev_desc = nullptr;
            return fail(VOX_ERR_OOM, std::string("runner descriptors: ") + hipGetErrorString(e));
        }
    }
    // stage through a pinned slot whose previous copy has completed
    const int k = r->slot;
    r->slot ^= 1;
    e = hipEventSynchronize(r->ev[k]);
    if (e != hipSuccess) return hip_fail(e, "hipEventSynchronize");
    if (plan.host.size() > r->pinned_cap[k]) {
        if (r->pinned[k]) (void)hipHostFree(r->pinned[k]);
        r->pinned_cap[k] = plan.host.size() * 2;
        e = hipHostMalloc(&r->pinned[k], r->pinned_cap[k], hipHostMallocDefault);
        if (e != hipSuccess) {
            r->pinned[k] = nullptr;
            r->pinned_cap[k] = 0;
            return fail(VOX_ERR_OOM, std::string("runner pinned staging: ") + hipGetErrorString(e));
        }
    }
    std::memcpy(r->pinned[k], plan.host.data(), plan.host.size());
    e = hipMemcpyAsync(r->dev_desc, r->pinned[k], plan.host.size(), hipMemcpyHostToDevice, r->stream);
    if (e == hipSuccess) e = hipEventRecord(r->ev[k], r->stream);
    if (e != hipSuccess) return hip_fail(e, "runner descriptor upload");
    bind_maps(&r->b, r->dev_desc);
    e = launch_all(&r->b, r->stream);
    if (e != hipSuccess) return hip_fail(e, "runner launch");
    r->launches++;
    return VOX_OK;
}

vox_status vox_runner_destroy(vox_runner* r) {
    if (!r) return VOX_OK;
    if (r->stream || r->b.ws || r->b.mws || r->dev_desc) (void)hipStreamSynchronize(r->stream);
    for (int i = 0; i < 2; i++) {
        if (r->ev[i]) (void)hipEventDestroy(r->ev[i]);
        if (r->pinned[i]) (void)hipHostFree(r->pinned[i]);
    }
    if (r->dev_desc) (void)hipFree(r->dev_desc);
    if (r->b.ws) (void)hipFree(r->b.ws);
    if (r->b.mws) (void)hipFree(r->b.mws);
    r->b.ws = nullptr;
    r->b.mws = nullptr;
    delete r;
    return VOX_OK;
}

}  // extern "C"
