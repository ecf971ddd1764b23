"""MI355X-native mesh voxelizer: Python host mirror of the reference API.

The reference (Forceflow/cuda_voxelizer v0.6) exposes, in C++:

* ``AABox<T>`` and ``voxinfo``                          (src/util.h:41-69)
* ``createMeshBBCube(AABox<float3>)``                   (src/util.h:80-110)
* ``voxelize(const voxinfo&, float*, unsigned int*, bool morton)``
                                                        (src/voxelize.cu:192-238)
* ``voxelize_solid(const voxinfo&, float*, unsigned int*, bool morton)``
                                                        (src/voxelize_solid.cu:147-193)

This module mirrors those names and argument meanings over the C-ABI of
``libvoxelizer_amd.so`` (hand-written HIP for gfx950).  Device buffers are
torch tensors on a HIP device; torch is used only for allocation and streams.
``voxelize``/``voxelize_solid`` keep the reference semantics: synchronous, and
OR/XOR into whatever the table holds.  ``voxelize_surface``/``voxelize_solid_table``
are the stream-ordered forms that define the whole table (fused zero-fill).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _capi
from ._capi import (ALGO_FLAGS, VOX_FLAG_ACCUMULATE, VOX_FLAG_ALGO_CHUNK, VOX_FLAG_ALGO_DIRECT, VOX_FLAG_ALGO_TILE,
                    VOX_LAYOUT_LINEAR, VOX_LAYOUT_MORTON, VoxError, check, vox_info)

__all__ = [
    "AABox", "VoxInfo", "create_mesh_bbox_cube", "createMeshBBCube", "voxelize", "voxelize_solid",
    "voxelize_surface", "voxelize_solid_table", "table_bytes", "table_words", "mesh_bbox",
    "voxinfo_for_mesh", "VoxError", "VOX_LAYOUT_LINEAR", "VOX_LAYOUT_MORTON", "VOX_FLAG_ACCUMULATE",
    "VOX_FLAG_ALGO_DIRECT", "VOX_FLAG_ALGO_TILE", "VOX_FLAG_ALGO_CHUNK", "ALGOS", "BatchJob", "Batch",
    "mesh_bbox_device", "pack_triangles", "upload", "convert_layout", "count_voxels", "check_voxel",
    "prepare_mesh", "sort_triangles",
]

ALGOS = tuple(ALGO_FLAGS)  # "auto", "direct", "tile", "chunk", "marker", "slab", "region"


def _algo_flag(algo: str) -> int:
    if algo not in ALGO_FLAGS:
        raise ValueError(f"algo must be one of {ALGOS}")
    return ALGO_FLAGS[algo]


@dataclass
class AABox:
    """Axis-aligned box (util.h:41-47); min/max are float32[3]."""
    min: np.ndarray
    max: np.ndarray

    def __post_init__(self):
        self.min = np.asarray(self.min, dtype=np.float32).reshape(3).copy()
        self.max = np.asarray(self.max, dtype=np.float32).reshape(3).copy()


class VoxInfo:
    """Mirror of the reference ``voxinfo`` (util.h:50-69).

    ``VoxInfo(bbox, gridsize, n_triangles)`` computes
    ``unit = (bbox.max - bbox.min) / float(gridsize)`` in binary32 (util.h:56-61)
    through the library, so host and kernels share one definition."""

    def __init__(self, bbox: AABox, gridsize, n_triangles: int):
        g = np.asarray(gridsize, dtype=np.int64).reshape(-1)
        if g.size == 3:
            if not (g[0] == g[1] == g[2]):
                raise ValueError("only cubic grids are supported (the reference builds cubic grids)")
            g = int(g[0])
        else:
            g = int(g[0])
        lib = _capi.load()
        self._c = vox_info()
        check(lib.vox_make_info(_f3(bbox.min), _f3(bbox.max), g, int(n_triangles), ctypes.byref(self._c)))

    @classmethod
    def _from_struct(cls, s: vox_info) -> "VoxInfo":
        obj = cls.__new__(cls)
        obj._c = s
        return obj

    @property
    def bbox(self) -> AABox:
        return AABox(np.array(self._c.bbox_min[:], np.float32), np.array(self._c.bbox_max[:], np.float32))

    @property
    def gridsize(self) -> tuple[int, int, int]:
        return tuple(int(v) for v in self._c.gridsize)

    @property
    def grid(self) -> int:
        return int(self._c.gridsize[0])

    @property
    def n_triangles(self) -> int:
        return int(self._c.n_triangles)

    @n_triangles.setter
    def n_triangles(self, n: int) -> None:
        self._c.n_triangles = int(n)

    @property
    def unit(self) -> np.ndarray:
        return np.array(self._c.unit[:], dtype=np.float32)

    def as_bytes(self) -> bytes:
        return ctypes.string_at(ctypes.addressof(self._c), 64)

    def __repr__(self) -> str:
        b = self.bbox
        return (f"VoxInfo(bbox=({b.min.tolist()}, {b.max.tolist()}), grid={self.gridsize}, "
                f"n_triangles={self.n_triangles}, unit={self.unit.tolist()})")


def _f3(a) -> ctypes.Array:
    a = np.asarray(a, dtype=np.float32).reshape(3)
    return (ctypes.c_float * 3)(*[float(x) for x in a])


def create_mesh_bbox_cube(box: AABox) -> AABox:
    """createMeshBBCube (util.h:80-110): pad the box to a cube, then pad every
    side by (max - min) / 10001."""
    lib = _capi.load()
    omin, omax = (ctypes.c_float * 3)(), (ctypes.c_float * 3)()
    check(lib.vox_bbox_cube(_f3(box.min), _f3(box.max), omin, omax))
    return AABox(np.array(omin[:], np.float32), np.array(omax[:], np.float32))


createMeshBBCube = create_mesh_bbox_cube


def table_bytes(grid: int) -> int:
    """Bytes of a grid^3 bit table: ceil(G^3/32)*4 (main.cpp:190)."""
    return int(_capi.load().vox_table_bytes(grid, grid, grid))


def table_words(grid: int) -> int:
    return table_bytes(grid) // 4


def mesh_bbox(vertices: np.ndarray) -> AABox:
    """Component-wise min/max over all vertices (trimesh2 ``need_bbox``, main.cpp:179)."""
    v = np.asarray(vertices, dtype=np.float32).reshape(-1, 3)
    return AABox(v.min(axis=0), v.max(axis=0))


def voxinfo_for_mesh(vertices: np.ndarray, grid: int, n_triangles: int) -> VoxInfo:
    """bbox -> createMeshBBCube -> voxinfo, as main.cpp:179-186 does."""
    return VoxInfo(create_mesh_bbox_cube(mesh_bbox(vertices)), (grid, grid, grid), n_triangles)


# ---- device entry points ------------------------------------------------------------
def _torch():
    import torch
    return torch


def _check_tensor(t, dtype_name: str, what: str):
    torch = _torch()
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise TypeError(f"{what} must be a torch tensor on a HIP device")
    if str(t.dtype) != dtype_name:
        raise TypeError(f"{what} must be {dtype_name}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{what} must be contiguous")


def _stream_ptr(stream) -> int | None:
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream) or None


class _OnStream:
    """Run a library call on `stream` (None: the current stream).

    A side stream first waits for the current stream (the inputs were written
    there), and tensors created inside the block are allocated on the side
    stream, so the caching allocator orders their reuse after its work.
    Results read on the host must be read after ``sync()``."""

    def __init__(self, stream, device=None):
        torch = _torch()
        self.cur = torch.cuda.current_stream(device)
        self.s = stream if stream is not None else self.cur
        self.side = self.s != self.cur
        self._ctx = torch.cuda.stream(self.s) if self.side else None

    def __enter__(self):
        if self.side:
            self.s.wait_stream(self.cur)
            self._ctx.__enter__()
        return self

    def __exit__(self, *exc):
        if self.side:
            self._ctx.__exit__(*exc)
        return False

    @property
    def ptr(self) -> int | None:
        return int(self.s.cuda_stream) or None

    def sync(self) -> None:
        self.s.synchronize()


def _info_checked(info: VoxInfo, tris, table, grid_words_needed=True):
    _check_tensor(tris, "torch.float32", "triangle_data")
    if tris.numel() < 9 * info.n_triangles:
        raise ValueError(f"triangle_data holds {tris.numel()} floats, need 9*{info.n_triangles}")
    if table is not None:
        if str(table.dtype) not in ("torch.int32", "torch.uint32"):
            raise TypeError("vtable must be int32/uint32")
        _check_tensor(table, str(table.dtype), "vtable")
        if table.numel() * 4 < table_bytes(info.grid):
            raise ValueError(f"vtable holds {table.numel()} words, need {table_words(info.grid)}")


def voxelize_surface(info: VoxInfo, tris, morton: bool = False, table=None, accumulate: bool = False,
                     algo: str = "auto", stream=None, tune: int = 0):
    """Surface voxelization (stream-ordered).  Returns the int32 table tensor.

    Without ``accumulate`` the table is fully defined by this call; with it
    bits are OR-ed into the existing contents (reference semantics).
    ``tune``: VOX_TUNE_* bits (``_capi.tune_*``) and the algorithm's VOX_SLAB_* / VOX_REGION_*
    bits (``_capi.region_*``); every setting gives the same table."""
    torch = _torch()
    with _OnStream(stream, tris.device) as on:
        if table is None:
            table = torch.empty(table_words(info.grid), dtype=torch.int32, device=tris.device)
        _info_checked(info, tris, table)
        flags = (VOX_FLAG_ACCUMULATE if accumulate else 0) | _algo_flag(algo) | (int(tune) & _capi.KNOB_MASK)
        lib = _capi.load()
        check(lib.vox_surface(ctypes.byref(info._c), tris.data_ptr(), table.data_ptr(),
                              VOX_LAYOUT_MORTON if morton else VOX_LAYOUT_LINEAR, flags, on.ptr))
    return table


def voxelize_solid_table(info: VoxInfo, tris, morton: bool = False, table=None, accumulate: bool = False,
                         algo: str = "auto", stream=None, offrow=None, tune: int = 0):
    """Solid (parity) voxelization (stream-ordered).  Returns the int32 table.

    ``offrow``: optional int64 device tensor of one element to which the
    number of column runs whose end left the row is added (the caller zeroes
    it).  Such runs spill into the next rows (linear layout) or are cut at
    the table end exactly as the reference indexes them; runs with a
    negative end are empty."""
    torch = _torch()
    with _OnStream(stream, tris.device) as on:
        if table is None:
            table = torch.empty(table_words(info.grid), dtype=torch.int32, device=tris.device)
        _info_checked(info, tris, table)
        flags = (VOX_FLAG_ACCUMULATE if accumulate else 0) | _algo_flag(algo) | (int(tune) & _capi.KNOB_MASK)
        cl_ptr = None
        if offrow is not None:
            _check_tensor(offrow, "torch.int64", "offrow")
            cl_ptr = offrow.data_ptr()
        lib = _capi.load()
        check(lib.vox_solid(ctypes.byref(info._c), tris.data_ptr(), table.data_ptr(),
                            VOX_LAYOUT_MORTON if morton else VOX_LAYOUT_LINEAR, flags, cl_ptr, on.ptr))
    return table


def voxelize(v: VoxInfo, triangle_data, vtable, use_thrust_path: bool = False, morton_code: bool = False):
    """Drop-in for ``voxelize(info, triangle_data, vtable, useThrustPath, morton_code)``.

    ORs the surface voxels into ``vtable`` (the caller zeroes it, as main.cpp:214/229
    do) and returns after the device is done, like voxelize.cu:218.
    ``use_thrust_path`` is accepted for source compatibility and has no effect."""
    torch = _torch()
    voxelize_surface(v, triangle_data, morton=morton_code, table=vtable, accumulate=True)
    torch.cuda.synchronize(triangle_data.device)


def voxelize_solid(v: VoxInfo, triangle_data, vtable, morton_code: bool = False):
    """Drop-in for ``voxelize_solid(info, triangle_data, vtable, morton_code)``:
    XORs the parity fill into ``vtable`` and synchronizes (voxelize_solid.cu:173)."""
    torch = _torch()
    voxelize_solid_table(v, triangle_data, morton=morton_code, table=vtable, accumulate=True)
    torch.cuda.synchronize(triangle_data.device)


# ---- batched mode -------------------------------------------------------------------
@dataclass
class BatchJob:
    """One voxelization inside a :class:`Batch`."""
    info: VoxInfo
    tris: object                 # float32 device tensor, 9 * n_triangles
    table: object                # int32 device tensor, table_words(G)
    morton: bool = False
    solid: bool = False
    accumulate: bool = False
    algo: str = "auto"
    offrow: object = None        # solid: optional int64 device tensor of 1 element
    tune: int = 0                # VOX_TUNE_* bits (autotuner knobs; the table is the same)


class Batch:
    """Many voxelizations replayed as one HIP graph launch (``vox_batch_*``).

    Pointers, voxinfo and flags are fixed when the batch is built; the
    contents of the triangle and table tensors may change between launches.
    The tensors must stay alive while the batch exists."""

    def __init__(self, jobs: list[BatchJob], use_graph: bool = True):
        lib = _capi.load()
        self._jobs = list(jobs)  # keep the tensors alive
        arr = (_capi.vox_batch_item * max(1, len(jobs)))()
        for i, j in enumerate(jobs):
            _info_checked(j.info, j.tris, j.table)
            it = arr[i]
            ctypes.memmove(ctypes.addressof(it.info), ctypes.addressof(j.info._c), 64)
            it.d_tris = j.tris.data_ptr()
            it.d_table = j.table.data_ptr()
            it.d_offrow = j.offrow.data_ptr() if j.offrow is not None else None
            it.layout = VOX_LAYOUT_MORTON if j.morton else VOX_LAYOUT_LINEAR
            it.solid = 1 if j.solid else 0
            it.flags = (VOX_FLAG_ACCUMULATE if j.accumulate else 0) | _algo_flag(j.algo) | (int(j.tune) & _capi.KNOB_MASK)
        handle = ctypes.c_void_p()
        check(lib.vox_batch_create(arr, len(jobs), 0 if use_graph else _capi.VOX_BATCH_NO_GRAPH,
                                   ctypes.byref(handle)))
        self._h = handle

    def launch(self, stream=None) -> None:
        check(_capi.load().vox_batch_launch(self._h, _stream_ptr(stream)))

    def info(self) -> dict:
        out = (ctypes.c_int64 * 8)()
        check(_capi.load().vox_batch_info(self._h, out))
        keys = ("jobs", "triangles", "chunk_jobs", "tile_surface_jobs", "tile_solid_jobs", "direct_jobs",
                "graph", "workspace_bytes")
        return dict(zip(keys, [int(v) for v in out]))

    def close(self) -> None:
        if getattr(self, "_h", None):
            _capi.load().vox_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- mesh preparation and table utilities -----------------------------------------------
def mesh_bbox_device(vertices, stream=None) -> AABox:
    """Device min/max of a float32 [V*3] vertex tensor (trimesh2 need_bbox)."""
    torch = _torch()
    _check_tensor(vertices, "torch.float32", "vertices")
    with _OnStream(stream, vertices.device) as on:
        out = torch.empty(6, dtype=torch.float32, device=vertices.device)
        check(_capi.load().vox_mesh_bbox(vertices.data_ptr(), vertices.numel() // 3, out.data_ptr(), on.ptr))
    on.sync()
    h = out.cpu().numpy()
    return AABox(h[:3], h[3:])


def pack_triangles(vertices, faces, stream=None):
    """Indexed mesh (float32 [V*3], int32 [F*3] device tensors) -> AoS float32
    [F*9] triangles (meshToGPU_managed).  Raises if a face index is out of range."""
    torch = _torch()
    _check_tensor(vertices, "torch.float32", "vertices")
    _check_tensor(faces, "torch.int32", "faces")
    nf = faces.numel() // 3
    with _OnStream(stream, vertices.device) as on:
        out = torch.empty(nf * 9, dtype=torch.float32, device=vertices.device)
        bad = torch.zeros(1, dtype=torch.int64, device=vertices.device)
        check(_capi.load().vox_pack_triangles(vertices.data_ptr(), vertices.numel() // 3, faces.data_ptr(), nf,
                                              out.data_ptr(), bad.data_ptr(), on.ptr))
    on.sync()  # the count is read here; `out` is complete for every stream after it
    if int(bad.item()):
        raise ValueError(f"{int(bad.item())} face indices out of range")
    return out


def sort_triangles(info: VoxInfo, tris, stream=None, with_perm: bool = False):
    """The triangles of `tris` (float32 [n*9] device tensor) in spatial order,
    for a mesh that is voxelized many times (vox_sort_triangles): a new tensor
    bucketed by the 16^3-voxel brick, at info's grid, of each triangle's lowest
    voxel corner, bricks in Morton order.  Every voxelization of it gives the
    same table as of `tris`, faster with the chunk algorithm when small
    consecutive input triangles lie far apart, slower for soups of large
    overlapping triangles.  With `with_perm`, also an int32 tensor:
    the index in `tris` of each output triangle."""
    torch = _torch()
    _check_tensor(tris, "torch.float32", "tris")
    n = int(info.n_triangles)
    if tris.numel() < 9 * n:
        raise ValueError(f"tris holds {tris.numel() // 9} triangles, info says {n}")
    with _OnStream(stream, tris.device) as on:
        out = torch.empty(9 * n, dtype=torch.float32, device=tris.device)
        perm = torch.empty(n, dtype=torch.int32, device=tris.device) if with_perm else None
        check(_capi.load().vox_sort_triangles(ctypes.byref(info._c), tris.data_ptr(), out.data_ptr(),
                                              perm.data_ptr() if with_perm else None, on.ptr))
    return (out, perm) if with_perm else out


def upload(host_array: np.ndarray, device="cuda:0", stream=None):
    """Host array -> new device tensor through pinned double-buffered staging."""
    torch = _torch()
    a = np.ascontiguousarray(host_array)
    out = torch.empty(a.shape, dtype=getattr(torch, str(a.dtype)), device=device)
    check(_capi.load().vox_upload_async(a.ctypes.data, a.nbytes, out.data_ptr(), _stream_ptr(stream)))
    torch.cuda.current_stream().synchronize() if stream is None else stream.synchronize()
    return out


def prepare_mesh(vertices: np.ndarray, faces: np.ndarray, grid: int, device="cuda:0", spatial_order: bool = False):
    """main.cpp:174-190 on the device: upload, bbox, createMeshBBCube, voxinfo,
    AoS packing.  Returns (VoxInfo, triangle tensor).  With `spatial_order`
    the triangles are then put in spatial order (sort_triangles), for a mesh
    that is voxelized many times."""
    v = upload(np.asarray(vertices, np.float32).reshape(-1), device)
    f = upload(np.asarray(faces, np.int32).reshape(-1), device)
    box = mesh_bbox_device(v)
    info = VoxInfo(create_mesh_bbox_cube(box), (grid, grid, grid), f.numel() // 3)
    tris = pack_triangles(v, f)
    return info, (sort_triangles(info, tris) if spatial_order else tris)


def convert_layout(table, grid: int, to_morton: bool, stream=None):
    """Linear <-> Morton re-layout of a power-of-two grid table (new tensor)."""
    torch = _torch()
    with _OnStream(stream, table.device) as on:
        out = torch.empty_like(table)
        check(_capi.load().vox_convert_layout(table.data_ptr(), out.data_ptr(), grid, int(to_morton), on.ptr))
    return out


def count_voxels(table, stream=None) -> int:
    torch = _torch()
    with _OnStream(stream, table.device) as on:
        c = torch.zeros(1, dtype=torch.int64, device=table.device)
        check(_capi.load().vox_count_voxels(table.data_ptr(), table.numel(), c.data_ptr(), on.ptr))
    on.sync()
    return int(c.item())


def check_voxel(host_table: np.ndarray, grid: int, x: int, y: int, z: int, morton: bool = False) -> bool:
    """checkVoxel (util.h:25-38) on a host table (either layout)."""
    t = np.ascontiguousarray(host_table, dtype=np.uint32)
    return bool(_capi.load().vox_check_voxel(t.ctypes.data, grid, x, y, z,
                                            VOX_LAYOUT_MORTON if morton else VOX_LAYOUT_LINEAR))
