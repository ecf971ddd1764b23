"""ctypes binding of libvoxelizer_amd.so (include/cuda_voxelizer_amd/vox_capi.h).

The library is built in-tree by ``__graft_entry__.build()`` (hipcc, gfx950).
Loading fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvoxelizer_amd.so")

VOX_OK = 0
VOX_ERR_INVALID = 1
VOX_ERR_HIP = 2
VOX_ERR_OOM = 3
VOX_ERR_UNSUPPORTED = 4
VOX_ERR_NO_DEVICE = 5

VOX_LAYOUT_LINEAR = 0
VOX_LAYOUT_MORTON = 1

VOX_FLAG_ACCUMULATE = 0x1
VOX_FLAG_ALGO_DIRECT = 0x2
VOX_FLAG_ALGO_TILE = 0x4
VOX_FLAG_ALGO_CHUNK = 0x8
VOX_FLAG_ALGO_MARKER = 0x10
VOX_FLAG_ALGO_SLAB = 0x20
VOX_FLAG_ALGO_REGION = 0x40
ALGO_FLAGS = {"auto": 0, "direct": VOX_FLAG_ALGO_DIRECT, "tile": VOX_FLAG_ALGO_TILE, "chunk": VOX_FLAG_ALGO_CHUNK,
              "marker": VOX_FLAG_ALGO_MARKER, "slab": VOX_FLAG_ALGO_SLAB, "region": VOX_FLAG_ALGO_REGION}



def region_words(lg: int) -> int:
    """VOX_REGION_WORDS: 2^lg table words per region (10..13)."""
    return ((lg - 9) & 15) << 8


def region_stage(lg: int) -> int:
    """VOX_REGION_STAGE: 2^lg staged pairs per chunk of triangles (0..10)."""
    return ((lg + 1) & 31) << 12


def region_desc(lg: int) -> int:
    """VOX_REGION_DESC: 2^lg descriptors per region (0..9)."""
    return ((lg + 1) & 15) << 17


def tune_big(lg: int) -> int:
    """VOX_TUNE_BIG: queue threshold 2^lg (chunk: voxels per box; marker: columns)."""
    return ((lg + 1) & 15) << 24


def tune_waves(w: int) -> int:
    """VOX_TUNE_WAVES: chunk kernel at 5 (w=1) or 6 (w=2) waves per SIMD."""
    return (w & 3) << 28


def tune_tile(t: int) -> int:
    """VOX_TUNE_TILE: tile algorithm at half (t=1) or twice (t=3) the built-in tile size."""
    return (t & 3) << 30


TUNE_MASK = 0xFF000000
ALGO_KNOB_MASK = 0x007FFF00  # VOX_SLAB_* / VOX_REGION_* bits (read only with their algorithm flag)
KNOB_MASK = TUNE_MASK | ALGO_KNOB_MASK


_STATUS_NAMES = {0: "VOX_OK", 1: "VOX_ERR_INVALID", 2: "VOX_ERR_HIP", 3: "VOX_ERR_OOM",
                 4: "VOX_ERR_UNSUPPORTED", 5: "VOX_ERR_NO_DEVICE"}


class VoxError(RuntimeError):
    """A C-ABI call returned a status other than VOX_OK."""

    def __init__(self, status: int, message: str):
        super().__init__(f"{_STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class vox_info(ctypes.Structure):
    """Byte layout of the reference ``voxinfo`` (util.h:50-61), 64 bytes."""
    _fields_ = [
        ("bbox_min", ctypes.c_float * 3),
        ("bbox_max", ctypes.c_float * 3),
        ("gridsize", ctypes.c_uint32 * 3),
        ("_pad0", ctypes.c_uint32),
        ("n_triangles", ctypes.c_uint64),
        ("unit", ctypes.c_float * 3),
        ("_pad1", ctypes.c_uint32),
    ]


assert ctypes.sizeof(vox_info) == 64


class vox_mesh(ctypes.Structure):
    _fields_ = [("vertices", ctypes.POINTER(ctypes.c_float)), ("n_vertices", ctypes.c_uint64),
                ("faces", ctypes.POINTER(ctypes.c_int32)), ("n_faces", ctypes.c_uint64)]


VOX_ENC_BINVOX = 1
VOX_ENC_OBJ_POINTS = 2
VOX_ENC_OBJ_CUBES = 3
VOX_ENC_VOX = 4


class vox_batch_item(ctypes.Structure):
    _fields_ = [
        ("info", vox_info),
        ("d_tris", ctypes.c_void_p),
        ("d_table", ctypes.c_void_p),
        ("d_offrow", ctypes.c_void_p),
        ("layout", ctypes.c_int32),
        ("solid", ctypes.c_int32),
        ("flags", ctypes.c_uint32),
        ("_pad", ctypes.c_uint32),
    ]


VOX_BATCH_NO_GRAPH = 0x1

_F3 = ctypes.c_float * 3
_P = ctypes.c_void_p

# name -> (restype, argtypes); every symbol declared in vox_capi.h
SIGNATURES = {
    "vox_init": (ctypes.c_int, [ctypes.c_int]),
    "vox_last_error": (ctypes.c_char_p, []),
    "vox_version": (ctypes.c_char_p, []),
    "vox_table_bytes": (ctypes.c_size_t, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]),
    "vox_make_info": (ctypes.c_int, [_F3, _F3, ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(vox_info)]),
    "vox_bbox_cube": (ctypes.c_int, [_F3, _F3, _F3, _F3]),
    "vox_surface": (ctypes.c_int, [ctypes.POINTER(vox_info), _P, _P, ctypes.c_int, ctypes.c_uint, _P]),
    "vox_solid": (ctypes.c_int, [ctypes.POINTER(vox_info), _P, _P, ctypes.c_int, ctypes.c_uint, _P, _P]),
    "vox_solid_noalloc": (ctypes.c_int, [ctypes.POINTER(vox_info), _P, _P, _P, _P]),
    "vox_noalloc_scratch_bytes": (ctypes.c_size_t, []),
    "vox_query_tile_plan": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int32)]),
    "vox_batch_create": (ctypes.c_int, [ctypes.POINTER(vox_batch_item), ctypes.c_int, ctypes.c_uint,
                                        ctypes.POINTER(ctypes.c_void_p)]),
    "vox_batch_launch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "vox_batch_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    "vox_batch_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "vox_mesh_bbox": (ctypes.c_int, [_P, ctypes.c_uint64, _P, _P]),
    "vox_pack_triangles": (ctypes.c_int, [_P, ctypes.c_uint64, _P, ctypes.c_uint64, _P, _P, _P]),
    "vox_upload_async": (ctypes.c_int, [_P, ctypes.c_size_t, _P, _P]),
    "vox_sort_triangles": (ctypes.c_int, [ctypes.POINTER(vox_info), _P, _P, _P, _P]),
    "vox_convert_layout": (ctypes.c_int, [_P, _P, ctypes.c_uint32, ctypes.c_int, _P]),
    "vox_count_voxels": (ctypes.c_int, [_P, ctypes.c_uint64, _P, _P]),
    "vox_check_voxel": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                       ctypes.c_int]),
    "vox_runner_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    "vox_runner_launch": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(vox_batch_item), ctypes.c_int]),
    "vox_runner_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "vox_encoder_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    "vox_encode": (ctypes.c_int, [ctypes.c_void_p, _P, ctypes.c_uint32, ctypes.c_int, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_uint64)]),
    "vox_encoder_data": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    "vox_encoder_copy": (ctypes.c_int, [ctypes.c_void_p, _P, ctypes.c_uint64]),
    "vox_encoder_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "vox_binvox_header": (ctypes.c_int, [ctypes.POINTER(vox_info), ctypes.c_char_p, ctypes.c_size_t]),
    "vox_load_mesh": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(vox_mesh)]),
    "vox_free_mesh": (None, [ctypes.POINTER(vox_mesh)]),
}

_lib = None


def load() -> ctypes.CDLL:
    """Load the HIP library (once) and bind every C-ABI symbol."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(hipcc, gfx950). There is no CPU fallback.")
    # One HIP runtime per process.  The torch wheel carries its own
    # libamdhip64 / libhsa-runtime64 under the same SONAMEs as /opt/rocm's but
    # asks for them by unversioned file name, so if this library (which needs
    # the versioned SONAMEs) were mapped first, importing torch would map a
    # second HIP and HSA runtime, and the one this library calls then finds
    # no device.  Importing torch first lets the dynamic linker bind this
    # library to the runtime torch already holds.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int) -> None:
    if status != VOX_OK:
        msg = load().vox_last_error()
        raise VoxError(status, msg.decode() if msg else "")
