/*
 * vox_capi.h -- C-ABI of the MI355X mesh voxelizer (libvoxelizer_amd.so).
 *
 * Plain pointers and sizes only; every entry point returns a vox_status and
 * never exits or throws.  Device pointers are HIP device (or managed)
 * allocations; `stream` is a hipStream_t passed as void* (NULL = the default
 * stream).  All compute entry points are stream-ordered and asynchronous with
 * respect to the host unless stated otherwise.
 *
 * Reference interfaces replaced (Forceflow/cuda_voxelizer v0.6, src/):
 *   vox_surface   <- void voxelize(const voxinfo&, float*, unsigned int*, bool)
 *                    voxelize.cu:192-238 (kernel voxelize_triangle :58-190)
 *   vox_solid     <- void voxelize_solid(const voxinfo&, float*, unsigned int*, bool)
 *                    voxelize_solid.cu:147-193 (kernel :73-145)
 *   vox_make_info <- voxinfo::voxinfo(AABox<float3>, uint3, size_t)  util.h:56-61
 *   vox_bbox_cube <- createMeshBBCube<float3>(AABox<float3>)         util.h:80-110
 *   vox_table_bytes <- vtable_size computation                       main.cpp:190
 *   vox_init      <- initCuda()                                      util_cuda.cpp:4-42
 */
#ifndef CUDA_VOXELIZER_AMD_VOX_CAPI_H
#define CUDA_VOXELIZER_AMD_VOX_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Byte-for-byte layout of the reference `voxinfo` (util.h:50-61): 64 bytes.
   offset 0 bbox.min, 12 bbox.max, 24 gridsize, 40 n_triangles, 48 unit. */
typedef struct vox_info {
    float bbox_min[3];
    float bbox_max[3];
    uint32_t gridsize[3];
    uint32_t _pad0;
    uint64_t n_triangles;
    float unit[3];
    uint32_t _pad1;
} vox_info;

typedef enum vox_status {
    VOX_OK = 0,
    VOX_ERR_INVALID = 1,     /* bad argument: see vox_last_error() */
    VOX_ERR_HIP = 2,         /* a HIP runtime call failed */
    VOX_ERR_OOM = 3,         /* device workspace allocation failed */
    VOX_ERR_UNSUPPORTED = 4, /* e.g. Morton layout with a non power-of-two grid */
    VOX_ERR_NO_DEVICE = 5    /* no HIP device visible */
} vox_status;

/* Output layouts of the bit table (MSB-first bits, word = idx >> 5,
   bit = 31 - (idx & 31); voxelize.cu:50-55). */
#define VOX_LAYOUT_LINEAR 0 /* idx = x + y*G + z*G*G   (voxelize.cu:175-178) */
#define VOX_LAYOUT_MORTON 1 /* idx = bit interleave x0 y0 z0 x1 ... (voxelize.cuh:20-34) */

/* Flags for the compute entry points. */
#define VOX_FLAG_ACCUMULATE 0x1u /* OR (surface) / XOR (solid) into the table's
                                    current contents, like the reference; without
                                    it the call defines every word of the table
                                    (the zero-fill is fused into the kernels). */
/* Algorithm selection (at most one; default: the library's choice per grid,
   layout and mode -- chunk for surface; for solid, tile when the mesh has fewer
   than G^2/4 triangles (large triangles), direct when it has at least 4 G^2 at
   G <= 64, marker otherwise).  All algorithms give identical tables. */
#define VOX_FLAG_ALGO_DIRECT 0x2u /* one triangle per lane, global atomics per word */
#define VOX_FLAG_ALGO_TILE 0x4u   /* triangles binned to output tiles; one workgroup
                                    builds each tile in LDS and writes it once */
#define VOX_FLAG_ALGO_CHUNK 0x8u  /* surface only: one workgroup per chunk of
                                    consecutive triangles merges words in an LDS
                                    hash, then one atomicOr per distinct word */
#define VOX_FLAG_ALGO_MARKER 0x10u /* solid only (G % 32 == 0, G <= 16384; Morton:
                                    power of two up to 4096; otherwise tile): one
                                    atomicXor marker per hit column at its run end
                                    in a zeroed bit table (the output itself for
                                    the linear layout), then every row becomes the
                                    suffix XOR of its markers along x */
#define VOX_FLAG_ALGO_SLAB 0x20u  /* linear layout, G % 32 == 0, G <= 2048: triangles
                                    binned to slabs of whole x rows; one workgroup
                                    builds each slab in LDS (surface: OR of the
                                    passing voxels; solid: run-end markers and the
                                    row suffix XOR) and writes every word of it
                                    once -- no zero-fill pass, no global atomics
                                    except for rare wide triangles and runs that
                                    leave their row */
#define VOX_FLAG_ALGO_REGION 0x40u /* power-of-two G in 32..2048 (otherwise the default):
                                    each workgroup merges its chunk's voxels (surface)
                                    or run-end markers (solid) in LDS and stages them
                                    as (word, mask) pairs per output region; then one
                                    workgroup per region merges its pairs in LDS (solid:
                                    and takes the row suffix XOR) and writes every word
                                    of the region once -- no zero-fill pass, and global
                                    atomics only for the staging reservations */
/* Region capacity bits (with VOX_FLAG_ALGO_REGION; a zero field is the library's
   choice).  Every setting gives the same table: pairs that find no room take a
   scratch-table fallback, so small capacities only cost time. */
#define VOX_REGION_WORDS(lg) ((((unsigned)(lg) - 9u) & 15u) << 8)  /* 2^lg words per region, 10..13 */
#define VOX_REGION_STAGE(lg) ((((unsigned)(lg) + 1u) & 31u) << 12) /* 2^lg staged pairs per chunk, 0..10 */
#define VOX_REGION_DESC(lg) ((((unsigned)(lg) + 1u) & 15u) << 17)  /* 2^lg descriptors per region, 0..9 */
/* Slab tuning bits (with VOX_FLAG_ALGO_SLAB; a zero field is the library's
   choice).  Every setting gives the same table. */
#define VOX_SLAB_TY(lg) ((((unsigned)(lg) + 1u) & 15u) << 8)     /* 2^lg rows in y per tile */
#define VOX_SLAB_TZ(lg) ((((unsigned)(lg) + 1u) & 15u) << 12)    /* 2^lg rows in z per tile */
#define VOX_SLAB_SPLIT(lg) ((((unsigned)(lg) + 1u) & 31u) << 16) /* surface: tiles with more than
                                    2^lg triangle entries are cut along x into 2, 4 or 8
                                    parts, one workgroup each (lg >= 30: never) */
/* Tuning bits for the default algorithms (autotuner knobs; a zero field is the
   built-in choice; every setting gives the same table):
   VOX_TUNE_BIG(lg)    chunk: voxel boxes above 2^lg voxels are shared by the workgroup
                       (built-in 256, plus boxes far larger than their wave's mean); marker: triangles
                       above 2^lg columns (built-in 32)
   VOX_TUNE_WAVES(w)   chunk: 1 = 5 waves per SIMD, 2 = 6 (built-in: 6 linear, 5 Morton)
   VOX_TUNE_TILE(t)    tile: 1 = half the built-in tile, 3 = twice (within the LDS limit) */
#define VOX_TUNE_BIG(lg) ((((unsigned)(lg) + 1u) & 15u) << 24)
#define VOX_TUNE_WAVES(w) ((((unsigned)(w)) & 3u) << 28)
#define VOX_TUNE_TILE(t) ((((unsigned)(t)) & 3u) << 30)
#define VOX_SLAB_WG(k) ((((unsigned)(k)) & 3u) << 21) /* raster workgroups: 1 = 1024 threads,
                                    2 = 256 threads (the default) */

/* ---- setup / diagnostics ---------------------------------------------------- */
vox_status vox_init(int device);            /* select + probe a device (initCuda) */
const char* vox_last_error(void);           /* thread-local message of the last failure */
const char* vox_version(void);

/* ceil(gx*gy*gz / 32) * 4, computed exactly in integers (main.cpp:190 uses a
   float division; identical for every grid whose voxel count is exact in fp32). */
size_t vox_table_bytes(uint32_t gx, uint32_t gy, uint32_t gz);

/* voxinfo constructor: unit = (max - min) / float(grid) per axis (util.h:56-61). */
vox_status vox_make_info(const float bbox_min[3], const float bbox_max[3], uint32_t grid,
                         uint64_t n_triangles, vox_info* out);

/* createMeshBBCube (util.h:80-110): pad to a cube, then by (max-min)/10001. */
vox_status vox_bbox_cube(const float in_min[3], const float in_max[3], float out_min[3],
                         float out_max[3]);

/* ---- hot path ------------------------------------------------------------------ */
/* Surface voxelization of info->n_triangles AoS FP32 triangles d_tris[9*n]
   (v0.xyz v1.xyz v2.xyz) into d_table (vox_table_bytes(G,G,G) bytes).
   Bit-exact with the reference CPU path (cpu_voxelizer.cpp:35-183). */
vox_status vox_surface(const vox_info* info, const float* d_tris, uint32_t* d_table, int layout,
                       unsigned flags, void* stream);

/* Solid (parity) voxelization; bit-exact with cpu_voxelizer.cpp:241-314.
   A hit column (y,z) XORs the bits x = 0..xmax, xmax = int(xc/unit.x - 0.5).
   When xmax leaves the row, the run follows the reference's indexing: for
   xmax >= G a linear run continues into the following rows (bits past the
   table end are dropped; a Morton run stops at G-1), and for xmax < 0 the
   run is empty (the reference CPU path would wrap its unsigned xmax).
   If d_offrow != NULL, the number of such runs is ADDED to *d_offrow
   (uint64 device memory, stream-ordered; the caller zeroes it). */
vox_status vox_solid(const vox_info* info, const float* d_tris, uint32_t* d_table, int layout,
                     unsigned flags, uint64_t* d_offrow, void* stream);

/* vox_solid without any device allocation.  It has no counterpart in the
   reference: it is for callers that cannot allocate device memory (code under
   HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
   Same table and *d_offrow as vox_solid(info, d_tris, d_table,
   VOX_LAYOUT_LINEAR, 0, d_offrow, stream), with the marker kernels; the
   table is fully defined by the call (no accumulation).  Limit: the grid edge
   must be a multiple of 32 from 32 to 16384 (VOX_ERR_UNSUPPORTED otherwise).
   Its scratch -- spill counters plus 65,536 out-of-row run records,
   vox_noalloc_scratch_bytes() = 1 MiB + 256 bytes whatever the input (more
   spilled runs are recomputed from the triangles) -- is a static device array
   of the library, one copy per GPU, held until the process ends; its address
   is taken in vox_init(device), or at the first call on a device.  Every
   call on one GPU shares that copy: calls that overlap on one GPU (different
   streams without ordering between them) race, so serialise them. */
vox_status vox_solid_noalloc(const vox_info* info, const float* d_tris, uint32_t* d_table,
                             uint64_t* d_offrow, void* stream);
size_t vox_noalloc_scratch_bytes(void);

/* ---- mesh preparation and table utilities (the steps either side of the path) --- */
/* Component-wise min/max of n_vertices xyz vertices (trimesh2 need_bbox, used at
   main.cpp:179), written to d_out6 = {min.xyz, max.xyz} (device memory). */
vox_status vox_mesh_bbox(const float* d_vertices, uint64_t n_vertices, float* d_out6, void* stream);
/* Indexed mesh -> AoS triangles d_tris[9*n_faces] (meshToGPU_managed,
   main.cpp:61-80).  Face indices outside [0, n_vertices) yield a zero vertex and
   are counted into *d_bad (added to; may be NULL). */
vox_status vox_pack_triangles(const float* d_vertices, uint64_t n_vertices, const int32_t* d_faces,
                              uint64_t n_faces, float* d_tris, uint64_t* d_bad, void* stream);
/* Spatial order for a mesh that is voxelized many times (no counterpart in the
   reference; for callers that keep meshes resident, such as serving loops).
   d_out[9*n] receives the triangles of d_tris (n = info->n_triangles) bucketed
   by the 16^3-voxel brick, at info's grid, of each triangle's lowest voxel
   corner, the bricks in Morton order; d_perm[n] (may be NULL) the index in
   d_tris of each output triangle.  Every voxelization of d_out, at any grid,
   layout and algorithm, gives the same table as of d_tris; the chunk
   algorithm merges more of its words in LDS and issues fewer atomics (1024^3
   torus of 1 M triangles 0.083 vs 0.107 ms per call, an icosphere whose faces
   interleave 0.126 vs 0.323 ms; the sort of the torus takes 0.036 ms).  It
   pays for surfaces of small triangles stored out of spatial order; it does
   not for soups of large overlapping triangles, whose sorted chunks pile
   their bits onto the same words (random soups: 1.9x to 3.4x slower).  Out of place: d_out and d_perm must not
   overlap d_tris or each other.  The order within a brick is unspecified.
   The library keeps per-stream counts and keys (4 bytes per triangle and up
   to 2 MiB), allocated at the first call or when a larger one comes. */
vox_status vox_sort_triangles(const vox_info* info, const float* d_tris, float* d_out, uint32_t* d_perm,
                              void* stream);
/* Host -> device copy through double-buffered pinned staging (returns once the
   last chunk is queued; the copies are ordered on `stream`). */
vox_status vox_upload_async(const void* host, size_t bytes, void* d_dst, void* stream);
/* Re-layout a power-of-two grid table: linear -> Morton (to_morton=1) or back. */
vox_status vox_convert_layout(const uint32_t* d_src, uint32_t* d_dst, uint32_t grid, int to_morton, void* stream);
/* Set voxels of a table of `words` words, added to *d_count (device). */
vox_status vox_count_voxels(const uint32_t* d_table, uint64_t words, uint64_t* d_count, void* stream);
/* checkVoxel (util.h:25-38) on a HOST copy of a table, either layout: 1 if set. */
int vox_check_voxel(const uint32_t* table, uint32_t grid, uint32_t x, uint32_t y, uint32_t z, int layout);

/* ---- batched mode (no counterpart in the reference) ---------------------------- */
/* Many voxelizations -- meshes of any size, grids, layouts, surface or solid --
   captured once into a HIP graph in which every kernel processes all jobs of
   its algorithm class; vox_batch_launch is then one graph launch.  Pointers,
   vox_info and flags are fixed at creation (a plan, as in FFT libraries);
   triangle and table contents may change between launches.  Tables of the
   items must not overlap. */
typedef struct vox_batch_item {
    vox_info info;         /* per-job voxinfo (cubic grid) */
    const float* d_tris;   /* device AoS triangles, info.n_triangles of them */
    uint32_t* d_table;     /* device table, vox_table_bytes(G,G,G) bytes */
    uint64_t* d_offrow;    /* solid only: out-of-row run counter (added to), or NULL */
    int32_t layout;        /* VOX_LAYOUT_LINEAR / VOX_LAYOUT_MORTON */
    int32_t solid;         /* 0 surface, 1 solid */
    uint32_t flags;        /* VOX_FLAG_ACCUMULATE, VOX_FLAG_ALGO_* (marker: linear,
                              non-accumulating jobs; others of its kind go to tile) */
    uint32_t _pad;
} vox_batch_item;

typedef struct vox_batch vox_batch;

#define VOX_BATCH_NO_GRAPH 0x1u /* launch the kernels directly each time (no graph) */

vox_status vox_batch_create(const vox_batch_item* items, int count, unsigned flags, vox_batch** out);
vox_status vox_batch_launch(vox_batch* batch, void* stream);
/* out[0..7] = jobs, triangles, chunk jobs, tile surface jobs, tile solid jobs,
   direct jobs, uses a graph, workspace bytes */
vox_status vox_batch_info(const vox_batch* batch, int64_t out[8]);
/* Waits for the device, then frees the batch. */
vox_status vox_batch_destroy(vox_batch* batch);

/* A re-fillable batch for serving: bound to one stream at creation; every
   vox_runner_launch plans the given jobs, stages their descriptors through
   pinned memory, copies them asynchronously and launches the batched kernels
   on that stream without waiting for the device (it waits only when a
   descriptor or workspace buffer has to grow).  Jobs of consecutive launches
   on one runner run in stream order; use several runners (streams) to overlap
   batches.  No graph: the job set changes every launch. */
typedef struct vox_runner vox_runner;
vox_status vox_runner_create(void* stream, vox_runner** out);
vox_status vox_runner_launch(vox_runner* runner, const vox_batch_item* items, int count);
vox_status vox_runner_destroy(vox_runner* runner);

/* ---- mesh files (the reference reads them with trimesh2, main.cpp:174-175) ------ */
/* Host arrays: vertices float[3*n_vertices], faces int32[3*n_faces]. */
typedef struct vox_mesh {
    float* vertices;
    uint64_t n_vertices;
    int32_t* faces;
    uint64_t n_faces;
} vox_mesh;
/* Loads .obj / .off / .stl / .ply (by extension) into *out (polygons split
   into fans around their first vertex; coordinates rounded once to binary32).
   Every face index is checked against the vertex count. */
vox_status vox_load_mesh(const char* path, vox_mesh* out);
void vox_free_mesh(vox_mesh* mesh);

/* ---- output encoders (the reference's writers, util_io.cpp) on the device ------- */
/* The reference writes its outputs on the host with one checkVoxel per voxel.
   An encoder permutes the table into the writer's traversal order, counts,
   scans and emits the file body on the device; the caller copies the bytes out.
     VOX_ENC_BINVOX      binvox RLE body: (value, count) byte pairs, y fastest,
                         then z, then x; a new pair after 255 equal voxels
                         (write_binvox, util_io.cpp:202-246; header below)
     VOX_ENC_OBJ_POINTS  "v x+0.5 y+0.5 z+0.5" per set voxel, z fastest, then y,
                         then x (write_obj_pointcloud, util_io.cpp:154-190)
     VOX_ENC_OBJ_CUBES   8 "v" + 12 "f" lines per set voxel in the same order
                         (write_cube / write_obj_cubes, util_io.cpp:45-133; the
                         reference then re-writes the file through trimesh2's
                         reorder_verts, which is not reproduced)
   Tables are cubic, either layout (Morton needs a power-of-two grid >= 4),
   grid edge 1..2048.  The morton output of the reference (write_binary of the
   raw table, util_io.cpp:192-200) needs no encoder: copy the table. */
#define VOX_ENC_BINVOX 1
#define VOX_ENC_OBJ_POINTS 2
#define VOX_ENC_OBJ_CUBES 3
#define VOX_ENC_VOX 4 /* complete MagicaVoxel .vox file (write_vox, util_io.cpp:249-285, through
                         the vendored VoxWriter: 126^3 cubes, XYZI chunks, scene graph, palette) */
typedef struct vox_encoder vox_encoder;
/* An encoder bound to a stream; it keeps grow-only device buffers. */
vox_status vox_encoder_create(void* stream, vox_encoder** out);
/* Encodes d_table (device memory) on the encoder's stream and waits for it
   once, to size the output; *out_bytes = length of the encoded body. */
vox_status vox_encode(vox_encoder* enc, const uint32_t* d_table, uint32_t grid, int layout, int format,
                      uint64_t* out_bytes);
/* Device pointer of the last encoded body (valid until the next vox_encode or
   vox_encoder_destroy on this encoder). */
vox_status vox_encoder_data(vox_encoder* enc, const uint8_t** d_bytes);
/* Copies the first `bytes` bytes of the last encoded body to host memory. */
vox_status vox_encoder_copy(vox_encoder* enc, void* host, uint64_t bytes);
vox_status vox_encoder_destroy(vox_encoder* enc);
/* The binvox ASCII header (util_io.cpp:211-216: dim, translate = bbox min,
   scale = largest bbox extent, %g formatting) into buf; returns its length as
   snprintf does, or -1 if info is NULL. */
int vox_binvox_header(const vox_info* info, char* buf, size_t cap);

/* ---- introspection ----------------------------------------------------------- */
/* Output tiling the tile algorithm uses for a grid/layout/mode (no counterpart
   in the reference).  out[0..7] = kind (0 linear rows, 1 Morton cubes, 2 Morton
   rows), tile x/y/z extent in voxels, number of tiles, LDS words per tile,
   words per row (row tiles), max tiles per binned triangle.  Returns
   VOX_ERR_UNSUPPORTED when the tile algorithm does not handle the case (the
   library then uses another algorithm). */
vox_status vox_query_tile_plan(uint32_t grid, int layout, int solid, int32_t out[8]);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* CUDA_VOXELIZER_AMD_VOX_CAPI_H */
