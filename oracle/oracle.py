"""ctypes wrapper of the CPU parity oracle (oracle/vox_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.  The oracle is a
plain-C restatement of the reference CPU voxelizer (src/cpu_voxelizer.cpp);
see the header of vox_oracle.c for the file:line map and for which parts are
pinned by which tests.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

MODE_REF = 0   # omp critical per bit (the reference's structure; the CPU baseline)
MODE_FAST = 1  # lock-free atomics per bit (same tables; used to check big cases)


class OracleInfo(ctypes.Structure):
    _fields_ = [("bmin", ctypes.c_float * 3), ("bmax", ctypes.c_float * 3), ("grid", ctypes.c_uint32 * 3),
                ("_pad0", ctypes.c_uint32), ("n_triangles", ctypes.c_uint64), ("unit", ctypes.c_float * 3),
                ("_pad1", ctypes.c_uint32)]


assert ctypes.sizeof(OracleInfo) == 64

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    lib.voxo_surface.restype = ctypes.c_double
    lib.voxo_surface.argtypes = [ctypes.POINTER(OracleInfo), P, ctypes.c_int64, P, P, ctypes.c_int, ctypes.c_int]
    lib.voxo_solid.restype = ctypes.c_double
    lib.voxo_solid.argtypes = [ctypes.POINTER(OracleInfo), P, ctypes.c_int64, P, P, ctypes.c_int, ctypes.c_int,
                               ctypes.POINTER(ctypes.c_int64)]
    lib.voxo_morton_encode.restype = ctypes.c_uint64
    lib.voxo_morton_encode.argtypes = [ctypes.c_uint32] * 3
    lib.voxo_morton_luts.restype = None
    lib.voxo_morton_luts.argtypes = [P]
    lib.voxo_mesh_bbox.restype = None
    lib.voxo_mesh_bbox.argtypes = [P, ctypes.c_int64, P, P]
    lib.voxo_bbox_cube.restype = None
    lib.voxo_bbox_cube.argtypes = [P, P, P, P]
    lib.voxo_make_info.restype = None
    lib.voxo_make_info.argtypes = [P, P, ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(OracleInfo)]
    lib.voxo_sizeof_info.restype = ctypes.c_int
    lib.voxo_set_threads.argtypes = [ctypes.c_int]
    lib.voxo_max_threads.restype = ctypes.c_int
    for fn in ("voxo_binvox_body", "voxo_obj_points", "voxo_obj_cubes", "voxo_vox_file"):
        getattr(lib, fn).restype = ctypes.c_uint64
        getattr(lib, fn).argtypes = [P, ctypes.c_uint32, ctypes.c_int, P, ctypes.c_uint64]
    _lib = lib
    return lib


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def info_from_bytes(b: bytes) -> OracleInfo:
    assert len(b) == 64
    return OracleInfo.from_buffer_copy(b)


def make_info(bmin, bmax, grid: int, n_triangles: int) -> OracleInfo:
    lib = load()
    bmin = np.ascontiguousarray(bmin, np.float32)
    bmax = np.ascontiguousarray(bmax, np.float32)
    out = OracleInfo()
    lib.voxo_make_info(_ptr(bmin), _ptr(bmax), grid, n_triangles, ctypes.byref(out))
    return out


def mesh_bbox(verts: np.ndarray):
    lib = load()
    v = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
    mn, mx = np.zeros(3, np.float32), np.zeros(3, np.float32)
    lib.voxo_mesh_bbox(_ptr(v), v.shape[0], _ptr(mn), _ptr(mx))
    return mn, mx


def bbox_cube(mn, mx):
    lib = load()
    mn = np.ascontiguousarray(mn, np.float32)
    mx = np.ascontiguousarray(mx, np.float32)
    omn, omx = np.zeros(3, np.float32), np.zeros(3, np.float32)
    lib.voxo_bbox_cube(_ptr(mn), _ptr(mx), _ptr(omn), _ptr(omx))
    return omn, omx


def info_for_mesh(verts: np.ndarray, grid: int, n_triangles: int) -> OracleInfo:
    """main.cpp:179-186: mesh bbox -> createMeshBBCube -> voxinfo."""
    mn, mx = mesh_bbox(verts)
    cmn, cmx = bbox_cube(mn, mx)
    return make_info(cmn, cmx, grid, n_triangles)


def morton_encode(x: int, y: int, z: int) -> int:
    return int(load().voxo_morton_encode(x, y, z))


def morton_luts() -> np.ndarray:
    out = np.zeros(768, np.uint32)
    load().voxo_morton_luts(_ptr(out))
    return out.reshape(3, 256)


def table_words(grid: int) -> int:
    return (grid ** 3 + 31) // 32


def _mesh_args(mesh):
    """mesh = AoS float32 [N,3,3] or (vertices [V,3], faces [F,3])."""
    if isinstance(mesh, tuple):
        verts, faces = mesh
        verts = np.array(verts, dtype=np.float32, copy=True).reshape(-1, 3)  # the prepass mutates it
        faces = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1, 3)
    else:
        tris = np.asarray(mesh, dtype=np.float32).reshape(-1, 3, 3)
        verts = np.array(tris.reshape(-1, 3), dtype=np.float32, copy=True)
        faces = np.arange(verts.shape[0], dtype=np.int32).reshape(-1, 3)
    return verts, faces


def surface(info: OracleInfo, mesh, morton: bool = False, mode: int = MODE_FAST, table=None,
            n_triangles: int | None = None):
    """Returns (table uint32[words], elapsed_ms).  ``table`` (optional) is
    OR-ed into, like the reference; otherwise a zeroed table is used."""
    lib = load()
    verts, faces = _mesh_args(mesh)
    info = OracleInfo.from_buffer_copy(bytes(info))
    info.n_triangles = faces.shape[0] if n_triangles is None else n_triangles
    if table is None:
        table = np.zeros(table_words(int(info.grid[0])), np.uint32)
    ms = lib.voxo_surface(ctypes.byref(info), _ptr(verts), verts.shape[0], _ptr(faces), _ptr(table),
                          int(morton), mode)
    return table, ms


def solid(info: OracleInfo, mesh, morton: bool = False, mode: int = MODE_FAST, table=None,
          n_triangles: int | None = None):
    """Returns (table uint32[words], offrow_runs, elapsed_ms); offrow_runs counts
    column runs whose end left the row (spilled or dropped, see vox_oracle.c)."""
    lib = load()
    verts, faces = _mesh_args(mesh)
    info = OracleInfo.from_buffer_copy(bytes(info))
    info.n_triangles = faces.shape[0] if n_triangles is None else n_triangles
    if table is None:
        table = np.zeros(table_words(int(info.grid[0])), np.uint32)
    offrow = ctypes.c_int64(0)
    ms = lib.voxo_solid(ctypes.byref(info), _ptr(verts), verts.shape[0], _ptr(faces), _ptr(table), int(morton),
                        mode, ctypes.byref(offrow))
    return table, int(offrow.value), ms


def set_threads(n: int) -> None:
    load().voxo_set_threads(n)


def max_threads() -> int:
    return int(load().voxo_max_threads())


def occupancy(table: np.ndarray, grid: int, morton: bool = False) -> np.ndarray:
    """Decode a bit table into a bool array [z, y, x] (MSB-first bits)."""
    n = grid ** 3
    bits = np.unpackbits(table.astype(">u4").view(np.uint8))[:n].astype(bool)
    if not morton:
        return bits.reshape(grid, grid, grid)
    z, y, x = np.meshgrid(np.arange(grid), np.arange(grid), np.arange(grid), indexing="ij")
    code = np.zeros(z.shape, dtype=np.uint64)
    for b in range(max(1, int(grid - 1).bit_length())):
        code |= ((x.astype(np.uint64) >> b) & 1) << (3 * b)
        code |= ((y.astype(np.uint64) >> b) & 1) << (3 * b + 1)
        code |= ((z.astype(np.uint64) >> b) & 1) << (3 * b + 2)
    return bits[code.astype(np.int64)]


def _writer(name: str, table: np.ndarray, grid: int, morton: bool) -> bytes:
    lib = load()
    t = np.ascontiguousarray(table).view(np.uint32)
    fn = getattr(lib, name)
    n = fn(_ptr(t), grid, 1 if morton else 0, None, 0)
    out = np.zeros(max(1, n), np.uint8)
    fn(_ptr(t), grid, 1 if morton else 0, _ptr(out), n)
    return out[:n].tobytes()


def binvox_body(table: np.ndarray, grid: int, morton: bool = False) -> bytes:
    """RLE body of write_binvox (util_io.cpp:218-245)."""
    return _writer("voxo_binvox_body", table, grid, morton)


def obj_points(table: np.ndarray, grid: int, morton: bool = False) -> bytes:
    """Text of write_obj_pointcloud (util_io.cpp:154-190)."""
    return _writer("voxo_obj_points", table, grid, morton)


def obj_cubes(table: np.ndarray, grid: int, morton: bool = False) -> bytes:
    """Text of write_obj_cubes before its trimesh2 re-write (util_io.cpp:45-133)."""
    return _writer("voxo_obj_cubes", table, grid, morton)


def vox_file(table: np.ndarray, grid: int, morton: bool = False) -> bytes:
    """MagicaVoxel file of write_vox (util_io.cpp:249-285, VoxWriter)."""
    return _writer("voxo_vox_file", table, grid, morton)
