"""Mesh files in, voxel files out.

Input mirrors the reference's trimesh2 read (main.cpp:174-175) through the
library's native loader (``vox_load_mesh``: OBJ, OFF, STL, PLY).  Output mirrors
the reference's writers (util_io.cpp) with the same file names and bodies; the
bodies are produced on the device by the encoders (``vox_encode``):

=====================  ==========================================  ==========================
reference writer       file name                                   body
=====================  ==========================================  ==========================
write_binary           ``<file>.bin``                              raw table (Morton output)
write_binvox           ``<file>_<G>.binvox``                       header + RLE pairs
write_obj_pointcloud   ``<file>_<G>_pointcloud.obj``               one ``v`` line per voxel
write_obj_cubes        ``<file>_<G>_voxels.obj``                   8 ``v`` + 12 ``f`` lines
write_vox              ``<file>_<G>.vox``                          MagicaVoxel chunks
=====================  ==========================================  ==========================

``<file>`` is the input path including its extension (the reference passes
``filename``, not ``filename_base``, main.cpp:244-258).  write_obj_cubes in
the reference finally re-reads and re-writes its file with trimesh2
(reorder_verts, util_io.cpp:139-149); that pass is not reproduced.  The
MagicaVoxel writer leaves palette entry 255 uninitialised; it is 0 here.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _capi
from ._capi import (VOX_ENC_BINVOX, VOX_ENC_OBJ_CUBES, VOX_ENC_OBJ_POINTS, VOX_ENC_VOX, VOX_LAYOUT_LINEAR,
                    VOX_LAYOUT_MORTON, check)

__all__ = ["load_mesh", "Encoder", "binvox_header", "encode", "write_binary", "write_binvox",
           "write_obj_pointcloud", "write_obj_cubes", "write_vox", "FORMATS"]

FORMATS = {"binvox": VOX_ENC_BINVOX, "obj_points": VOX_ENC_OBJ_POINTS, "obj": VOX_ENC_OBJ_CUBES, "vox": VOX_ENC_VOX}


def load_mesh(path: str):
    """(vertices float32 [nv, 3], faces int32 [nf, 3]) of a mesh file."""
    lib = _capi.load()
    m = _capi.vox_mesh()
    check(lib.vox_load_mesh(str(path).encode(), ctypes.byref(m)))
    try:
        nv, nf = int(m.n_vertices), int(m.n_faces)
        v = np.ctypeslib.as_array(m.vertices, shape=(max(1, nv * 3),))[: nv * 3].copy().reshape(nv, 3)
        f = np.ctypeslib.as_array(m.faces, shape=(max(1, nf * 3),))[: nf * 3].copy().reshape(nf, 3)
    finally:
        lib.vox_free_mesh(ctypes.byref(m))
    return v, f


def _device_table(table, device="cuda"):
    import torch
    if isinstance(table, np.ndarray):
        table = torch.from_numpy(np.ascontiguousarray(table).view(np.int32))
    if not table.is_cuda:
        table = table.to(device)
    return table.contiguous()


class Encoder:
    """A device encoder bound to a stream (``vox_encoder_*``)."""

    def __init__(self, stream=None):
        import torch
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        h = ctypes.c_void_p()
        check(_capi.load().vox_encoder_create(ctypes.c_void_p(self.stream.cuda_stream), ctypes.byref(h)))
        self.h = h

    def encode(self, table, grid: int, morton: bool, fmt: int) -> bytes:
        t = _device_table(table)
        n = ctypes.c_uint64()
        lib = _capi.load()
        check(lib.vox_encode(self.h, ctypes.c_void_p(t.data_ptr()), grid,
                             VOX_LAYOUT_MORTON if morton else VOX_LAYOUT_LINEAR, fmt, ctypes.byref(n)))
        out = np.empty(max(1, n.value), np.uint8)
        check(lib.vox_encoder_copy(self.h, ctypes.c_void_p(out.ctypes.data), n.value))
        return out[: n.value].tobytes()

    def close(self) -> None:
        if self.h:
            check(_capi.load().vox_encoder_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def encode(table, grid: int, fmt: str = "binvox", morton: bool = False) -> bytes:
    """Encoded output ("binvox" body, "obj_points", "obj", or a whole "vox" file) of a table."""
    with Encoder() as enc:
        return enc.encode(table, grid, morton, FORMATS[fmt])


def binvox_header(info) -> bytes:
    buf = ctypes.create_string_buffer(512)
    n = _capi.load().vox_binvox_header(ctypes.byref(info._c), buf, 512)
    if n < 0 or n >= 512:
        raise ValueError("binvox header did not fit")
    return buf.raw[:n]


def write_binary(table, filename: str) -> str:
    """write_binary (util_io.cpp:192-200): the raw table to ``<filename>.bin``."""
    import torch
    t = table.cpu() if isinstance(table, torch.Tensor) else torch.from_numpy(np.asarray(table))
    path = filename + ".bin"
    with open(path, "wb") as fh:
        fh.write(t.numpy().tobytes())
    return path


def _write(path: str, head: bytes, body: bytes) -> str:
    with open(path, "wb") as fh:
        fh.write(head)
        fh.write(body)
    return path


def write_binvox(table, info, filename: str, morton: bool = False) -> str:
    """write_binvox (util_io.cpp:202-246) -> ``<filename>_<G>.binvox``."""
    G = info.grid
    return _write(f"{filename}_{G}.binvox", binvox_header(info), encode(table, G, "binvox", morton))


def write_obj_pointcloud(table, info, filename: str, morton: bool = False) -> str:
    """write_obj_pointcloud (util_io.cpp:154-190) -> ``<filename>_<G>_pointcloud.obj``."""
    G = info.grid
    return _write(f"{filename}_{G}_pointcloud.obj", b"", encode(table, G, "obj_points", morton))


def write_obj_cubes(table, info, filename: str, morton: bool = False) -> str:
    """write_obj_cubes (util_io.cpp:92-152, without the trimesh2 re-write) -> ``<filename>_<G>_voxels.obj``."""
    G = info.grid
    return _write(f"{filename}_{G}_voxels.obj", b"", encode(table, G, "obj", morton))


def write_vox(table, info, filename: str, morton: bool = False) -> str:
    """write_vox (util_io.cpp:249-285, VoxWriter) -> ``<filename>_<G>.vox``."""
    G = info.grid
    return _write(f"{filename}_{G}.vox", b"", encode(table, G, "vox", morton))
