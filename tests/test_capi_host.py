"""CPU tests of libvoxelizer_amd.so: it loads, exports every symbol its
headers declare, and its host-side rules (voxinfo, createMeshBBCube, table
size, argument validation) agree with the oracle and the reference's closed
forms.  No compute entry point is called with valid arguments here."""
import ctypes
import os
import re

import numpy as np
import pytest

from tests.conftest import ROOT

HEADERS = [os.path.join(ROOT, "include", "cuda_voxelizer_amd", h) for h in sorted(
    os.listdir(os.path.join(ROOT, "include", "cuda_voxelizer_amd"))) if h.endswith(".h")]


def _declared_c_symbols():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        if 'extern "C"' not in text:
            continue
        for m in re.finditer(r"^\s*(?:const\s+)?[\w\s\*]+?\b(vox_\w+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol(capi):
    names = _declared_c_symbols()
    assert len(names) >= 8
    missing = [n for n in names if not hasattr(capi, n)]
    assert not missing, missing


def test_cpp_dropin_symbols_exported(capi):
    out = os.popen(f"nm -D --defined-only {capi._name} | c++filt").read()
    for sig in ("voxelize(voxinfo const&, float*, unsigned int*, bool)",
                "voxelize(voxinfo const&, float*, unsigned int*, bool, bool)",
                "voxelize_solid(voxinfo const&, float*, unsigned int*, bool)",
                "createMeshBBCube(AABox<HIP_vector_type<float, 3u> >)"):
        assert sig in out, sig


def test_python_binding_covers_header(capi):
    from cuda_voxelizer_amd import _capi
    assert _declared_c_symbols() <= set(_capi.SIGNATURES)


def test_make_info_and_bbox_cube_match_oracle(oracle):
    import cuda_voxelizer_amd as V
    rng = np.random.default_rng(0)
    for g in (1, 7, 64, 100, 512, 1024):
        v = rng.normal(size=(500, 3)).astype(np.float32) * rng.uniform(0.1, 5, 3).astype(np.float32)
        info = V.voxinfo_for_mesh(v, g, 123)
        oinfo = oracle.info_for_mesh(v, g, 123)
        assert info.as_bytes() == bytes(oinfo)
    box = V.create_mesh_bbox_cube(V.AABox([1, 2, 3], [4, 4, 4]))
    omn, omx = oracle.bbox_cube([1, 2, 3], [4, 4, 4])
    np.testing.assert_array_equal(box.min, omn)
    np.testing.assert_array_equal(box.max, omx)


def test_voxinfo_layout_is_reference_layout():
    from cuda_voxelizer_amd._capi import vox_info
    assert ctypes.sizeof(vox_info) == 64
    assert vox_info.gridsize.offset == 24 and vox_info.n_triangles.offset == 40 and vox_info.unit.offset == 48


@pytest.mark.parametrize("g", [1, 2, 3, 7, 32, 33, 100, 512, 1000, 1024, 2048])
def test_table_bytes(capi, g):
    assert capi.vox_table_bytes(g, g, g) == ((g ** 3 + 31) // 32) * 4


def test_compute_entry_points_validate_before_touching_the_device(capi):
    from cuda_voxelizer_amd import _capi
    import cuda_voxelizer_amd as V
    info = V.VoxInfo(V.AABox([0, 0, 0], [1, 1, 1]), (8, 8, 8), 1)
    dummy = ctypes.c_void_p(0x1000)  # never dereferenced: validation fails first
    assert capi.vox_surface(None, dummy, dummy, 0, 0, None) == _capi.VOX_ERR_INVALID
    assert b"NULL" in capi.vox_last_error()
    bad = _capi.vox_info.from_buffer_copy(info.as_bytes())
    bad.gridsize[2] = 4
    assert capi.vox_surface(ctypes.byref(bad), dummy, dummy, 0, 0, None) == _capi.VOX_ERR_INVALID
    assert b"cubic" in capi.vox_last_error()
    npot = V.VoxInfo(V.AABox([0, 0, 0], [1, 1, 1]), (12, 12, 12), 1)
    assert capi.vox_surface(ctypes.byref(npot._c), dummy, dummy, 1, 0, None) == _capi.VOX_ERR_UNSUPPORTED
    assert capi.vox_solid(ctypes.byref(npot._c), dummy, dummy, 1, 0, None, None) == _capi.VOX_ERR_UNSUPPORTED
    assert capi.vox_surface(ctypes.byref(info._c), dummy, dummy, 7, 0, None) == _capi.VOX_ERR_INVALID
    assert capi.vox_surface(ctypes.byref(info._c), None, dummy, 0, 0, None) == _capi.VOX_ERR_INVALID
    zero = V.VoxInfo(V.AABox([0, 0, 0], [0, 0, 0]), (8, 8, 8), 1)
    assert capi.vox_surface(ctypes.byref(zero._c), dummy, dummy, 0, 0, None) == _capi.VOX_ERR_INVALID
    with pytest.raises(V.VoxError):
        V._capi.check(capi.vox_surface(None, dummy, dummy, 0, 0, None))
    two = _capi.VOX_FLAG_ALGO_TILE | _capi.VOX_FLAG_ALGO_CHUNK
    assert capi.vox_surface(ctypes.byref(info._c), dummy, dummy, 0, two, None) == _capi.VOX_ERR_INVALID
    assert capi.vox_solid(ctypes.byref(info._c), dummy, dummy, 0, _capi.VOX_FLAG_ALGO_CHUNK, None,
                          None) == _capi.VOX_ERR_INVALID


def test_tile_plans_respect_kernel_bounds(capi):
    """Every tile plan the library would use fits the kernels' assumptions:
    tiles partition the grid, <= 65536 tiles (the scan), <= 64 KiB LDS per tile,
    rows of <= 64 words (the suffix-XOR wave), Morton tiles are aligned cubes."""
    out = (ctypes.c_int32 * 8)()
    seen = 0
    for g in list(range(1, 130)) + [192, 256, 384, 512, 640, 1024, 1536, 2048, 4096]:
        for layout in (0, 1):
            for solid in (0, 1):
                st = capi.vox_query_tile_plan(g, layout, solid, out)
                if st != 0:
                    continue
                seen += 1
                kind, tx, ty, tz, ntiles, words, roww, kmax = list(out)
                assert g % tx == 0 and g % ty == 0 and g % tz == 0
                assert ntiles == (g // tx) * (g // ty) * (g // tz) <= 65536
                assert 0 < words * 4 <= 65536
                assert kmax >= 1
                if kind == 0:
                    assert layout == 0 and g % 32 == 0 and tx == g and roww == g // 32 <= 64
                    assert words == roww * ty * tz
                elif kind == 1:
                    assert layout == 1 and not solid and tx == ty == tz and tx <= 32 and tx * tx * tx >= 64
                    assert (g & (g - 1)) == 0 and words * 32 == tx ** 3
                else:
                    assert layout == 1 and solid and tx == g and ty == tz and ty ** 3 >= 32 and roww == g // 32 <= 64
                    assert words == roww * ty * tz
    assert seen > 20


def test_check_voxel_matches_layout_decoding(oracle):
    import cuda_voxelizer_amd as V
    from cuda_voxelizer_amd import meshgen
    v, f = meshgen.icosphere_jitter(level=2, seed=3)
    info = oracle.info_for_mesh(v, 16, f.shape[0])
    for morton in (False, True):
        t, _ = oracle.surface(info, (v, f), morton=morton)
        occ = oracle.occupancy(t, 16, morton=morton)
        for (z, y, x) in [(0, 0, 0), (15, 15, 15)] + [tuple(p) for p in np.argwhere(occ)[:50]]:
            assert V.check_voxel(t, 16, int(x), int(y), int(z), morton) == bool(occ[z, y, x])
        assert not V.check_voxel(t, 16, 16, 0, 0, morton)  # out of range reads nothing
