"""GPU test of the C++ drop-in from C++: tests/cpp/dropin_caller.cpp follows
main.cpp:174-221 against <cuda_voxelizer_amd/voxelize.h> (need_bbox,
createMeshBBCube<float3>, voxinfo, the float vtable size, triangle packing,
the 4- and 5-argument voxelize and voxelize_solid) and the tables it gets
equal the oracle's, as does the voxinfo it builds."""
import ctypes
import os

import numpy as np
import pytest

from tests.conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu

LIB = os.path.join(ROOT, "tests", "cpp", "libdropin_caller.so")


@pytest.fixture(scope="module")
def dropin(gpu):
    from cuda_voxelizer_amd import _capi
    _capi.load()  # the HIP runtime torch uses, before the helper maps the library
    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} is missing: build it with __graft_entry__.build()")
    lib = ctypes.CDLL(LIB)
    P = ctypes.c_void_p
    lib.dropin_run.argtypes = [P, ctypes.c_int, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int, P, P]
    lib.dropin_run.restype = ctypes.c_int
    return lib


def _meshes():
    from cuda_voxelizer_amd import meshgen
    v, f = meshgen.icosphere_jitter()  # the 64^3 workload: 5,120 triangles
    g = np.load(os.path.join(GOLDEN, "bunny.npz"))  # the reference's test model
    return {"icosphere": (np.ascontiguousarray(v, np.float32), np.ascontiguousarray(f, np.int32)),
            "bunny": (np.ascontiguousarray(g["vertices"], np.float32), np.ascontiguousarray(g["faces"], np.int32))}


@pytest.mark.parametrize("mesh", ["icosphere", "bunny"])
@pytest.mark.parametrize("grid", [64, 128])
def test_cpp_dropin_matches_oracle(dropin, oracle, mesh, grid):
    v, f = _meshes()[mesh]
    words = (grid ** 3 + 31) // 32
    for solid, morton, five in [(0, 0, 0), (0, 1, 1), (0, 0, 1), (1, 0, 0), (1, 1, 0)]:
        out = np.zeros(words, np.uint32)
        info = np.zeros(16, np.float32)  # the 64-byte voxinfo the caller built
        rc = dropin.dropin_run(v.ctypes.data, v.shape[0], f.ctypes.data, f.shape[0], grid, solid, morton, five,
                               out.ctypes.data, info.ctypes.data)
        assert rc == 0
        oinfo = oracle.info_for_mesh(v, grid, f.shape[0])
        # voxinfo: bbox.min, bbox.max, gridsize (uint3), n_triangles (size_t), unit (padding not compared)
        raw = info.view(np.uint8)
        assert raw[0:24].tobytes() == bytes(oinfo)[0:24]
        assert raw[24:36].tobytes() == bytes(oinfo)[24:36]
        assert raw[40:48].tobytes() == bytes(oinfo)[40:48]
        assert raw[48:60].tobytes() == bytes(oinfo)[48:60]
        if solid:
            exp, _, _ = oracle.solid(oinfo, (v, f), morton=bool(morton))
        else:
            exp, _ = oracle.surface(oinfo, (v, f), morton=bool(morton))
        assert (out == exp).all(), (mesh, grid, solid, morton, five)
