"""GPU tests of the steps either side of the path: device bbox and packing
against the host definitions, layout conversion round trips, voxel counts."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_prepare_mesh_matches_host_prep(oracle, gpu):
    import cuda_voxelizer_amd as V
    from cuda_voxelizer_amd import meshgen
    v, f = meshgen.icosphere_sh(level=5, seed=5)
    v = v + np.float32(-0.37)  # negative coordinates exercise the ordered-key min/max
    info, tris = V.prepare_mesh(v, f, 128)
    host_info = V.voxinfo_for_mesh(v, 128, f.shape[0])
    assert info.as_bytes() == host_info.as_bytes()
    np.testing.assert_array_equal(tris.cpu().numpy().reshape(-1, 3, 3), meshgen.to_aos(v, f))
    got = V.voxelize_surface(info, tris).cpu().numpy().view(np.uint32)
    exp, _ = oracle.surface(oracle.info_from_bytes(info.as_bytes()), (v, f))
    assert (got == exp).all()


def test_pack_rejects_bad_indices(gpu):
    import torch
    import cuda_voxelizer_amd as V
    vtx = torch.zeros(3 * 4, dtype=torch.float32, device=gpu)
    faces = torch.tensor([0, 1, 2, 1, 2, 7], dtype=torch.int32, device=gpu)
    with pytest.raises(ValueError):
        V.pack_triangles(vtx, faces)


@pytest.mark.parametrize("grid", [4, 16, 32, 64, 128, 256, 1024, 2048])
def test_layout_conversion_round_trip(oracle, gpu, grid):
    import torch
    import cuda_voxelizer_amd as V
    from cuda_voxelizer_amd import meshgen
    v, f = meshgen.icosphere_jitter(level=3, seed=grid)
    info = V.voxinfo_for_mesh(v, grid, f.shape[0])
    d = torch.from_numpy(meshgen.to_aos(v, f).reshape(-1)).to(gpu)
    lin = V.voxelize_solid_table(info, d)
    mor = V.voxelize_solid_table(info, d, morton=True)
    assert (V.convert_layout(lin, grid, True) == mor).all()
    assert (V.convert_layout(mor, grid, False) == lin).all()
    n = V.count_voxels(lin)
    assert n == V.count_voxels(mor) == int(np.unpackbits(lin.cpu().numpy().view(np.uint8)).sum())


def _morton_codes(g):
    c = np.arange(g, dtype=np.uint64)
    sp = np.zeros(g, np.uint64)
    for k in range(12):
        sp |= ((c >> np.uint64(k)) & np.uint64(1)) << np.uint64(3 * k)
    # [z, y, x] -> code of (x, y, z): x bits at 3k, y at 3k+1, z at 3k+2
    return (sp[None, None, :] | (sp[None, :, None] << np.uint64(1)) | (sp[:, None, None] << np.uint64(2))).reshape(-1)


@pytest.mark.parametrize("grid", [8, 32, 64, 256])
def test_layout_conversion_random_bits(gpu, grid):
    """Random tables both ways against a numpy restatement of the two layouts
    (linear x + y G + z G^2, Morton bit interleave, MSB-first words)."""
    import torch
    import cuda_voxelizer_amd as V
    rng = np.random.default_rng(grid)
    words = grid ** 3 // 32
    lin = rng.integers(0, 2 ** 32, size=words, dtype=np.uint64).astype(np.uint32)
    bits = np.unpackbits(lin.byteswap().view(np.uint8))  # voxel order, MSB first
    code = _morton_codes(grid)
    mbits = np.zeros_like(bits)
    mbits[code] = bits
    mor = np.packbits(mbits).view(np.uint32).byteswap()
    d_lin = torch.from_numpy(lin.view(np.int32)).to(gpu)
    d_mor = torch.from_numpy(mor.view(np.int32)).to(gpu)
    assert np.array_equal(V.convert_layout(d_lin, grid, True).cpu().numpy().view(np.uint32), mor)
    assert np.array_equal(V.convert_layout(d_mor, grid, False).cpu().numpy().view(np.uint32), lin)


@pytest.mark.parametrize("nv,nf", [(3, 1), (777, 1000), (70001, 300001)])
def test_bbox_and_pack_odd_sizes(gpu, nv, nf):
    """Partial groups of faces and vertices: device bbox and packing equal numpy."""
    import torch
    import cuda_voxelizer_amd as V
    rng = np.random.default_rng(nv)
    v = rng.normal(size=(nv, 3)).astype(np.float32) * np.float32(3.0) - np.float32(1.0)
    f = rng.integers(0, nv, size=(nf, 3)).astype(np.int32)
    dv = torch.from_numpy(v.reshape(-1)).to(gpu)
    df = torch.from_numpy(f.reshape(-1)).to(gpu)
    box = V.mesh_bbox_device(dv)
    np.testing.assert_array_equal(box.min, v.min(0))
    np.testing.assert_array_equal(box.max, v.max(0))
    tris = V.pack_triangles(dv, df)
    np.testing.assert_array_equal(tris.cpu().numpy(), v[f].reshape(-1))


def test_side_stream_calls_are_ordered(oracle, gpu):
    """Every helper that takes a stream runs on it, after the caller's work on
    the current stream, and returns results that stream produced: inputs are
    written on the current stream immediately before each call."""
    import torch
    import cuda_voxelizer_amd as V
    from cuda_voxelizer_amd import meshgen
    side = torch.cuda.Stream()
    v, f = meshgen.icosphere_sh(level=6, seed=11)
    v = v + np.float32(-0.5)
    dv = torch.empty(v.size, dtype=torch.float32, device=gpu)
    df = torch.empty(f.size, dtype=torch.int32, device=gpu)
    for rep in range(3):
        dv.fill_(7.0)
        df.fill_(0)
        dv.copy_(torch.from_numpy(v.reshape(-1)).to(gpu))  # written on the current stream
        df.copy_(torch.from_numpy(f.reshape(-1)).to(gpu))
        box = V.mesh_bbox_device(dv, stream=side)
        np.testing.assert_array_equal(box.min, v.min(0))
        np.testing.assert_array_equal(box.max, v.max(0))
        tris = V.pack_triangles(dv, df, stream=side)
        np.testing.assert_array_equal(tris.cpu().numpy(), meshgen.to_aos(v, f).reshape(-1))
        info = V.voxinfo_for_mesh(v, 128, f.shape[0])
        table = V.voxelize_surface(info, tris, stream=side)
        side.synchronize()
        exp, _ = oracle.surface(oracle.info_from_bytes(info.as_bytes()), (v, f))
        assert (table.cpu().numpy().view(np.uint32) == exp).all()
        assert V.count_voxels(table, stream=side) == int(np.unpackbits(exp.view(np.uint8)).sum())
        mor = V.convert_layout(table, 128, True, stream=side)
        side.synchronize()
        exp_m, _ = oracle.surface(oracle.info_from_bytes(info.as_bytes()), (v, f), morton=True)
        assert (mor.cpu().numpy().view(np.uint32) == exp_m).all()


# ---- spatial order (vox_sort_triangles) ---------------------------------------------
def _brick_keys(tris, info):
    """Host restatement of the sort key: the voxel box's lowest corner at info's
    grid (the kernels' binary32 arithmetic: v - bbox_min, min over the
    vertices, / unit, truncated, clamped), in bricks of 16 voxels (2^(lg G - 6)
    above 1024), Morton-interleaved."""
    G = info.grid
    bmin = np.asarray(info.bbox.min, np.float32)
    unit = info.unit.astype(np.float32)
    lo = ((tris.reshape(-1, 3, 3) - bmin).min(axis=1) / unit)
    lo = np.clip(np.trunc(lo), 0, G - 1).astype(np.int64)
    lg = int(np.ceil(np.log2(G))) if G > 1 else 0
    shift = max(4, lg - 6)
    b = lo >> shift
    key = np.zeros(len(b), np.int64)
    for bit in range(7):
        for ax in range(3):
            key |= ((b[:, ax] >> bit) & 1) << (3 * bit + ax)
    return key


def _check_sorted(oracle, gpu, tris, grid, layouts=(False,), solid=(False, True), algos=("auto", "chunk")):
    import torch
    import cuda_voxelizer_amd as V
    tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 3, 3)
    n = tris.shape[0]
    info = V.voxinfo_for_mesh(tris.reshape(-1, 3), grid, n)
    d = torch.from_numpy(tris.reshape(-1)).to(gpu)
    print(f"sort case: n={n} grid={grid}", flush=True)
    out, perm = V.sort_triangles(info, d, with_perm=True)
    p = perm.cpu().numpy().astype(np.int64)
    o = out.cpu().numpy().reshape(-1, 3, 3)
    assert (np.sort(p) == np.arange(n)).all()  # a permutation
    assert (o.view(np.uint32) == tris[p].view(np.uint32)).all()  # the same triangles, bit for bit
    k = _brick_keys(o, info)
    assert (np.diff(k) >= 0).all()  # bricks in Morton order
    oi = oracle.info_from_bytes(info.as_bytes())
    for morton in layouts:
        for s in solid:
            exp = oracle.solid(oi, tris, morton=morton)[0] if s else oracle.surface(oi, tris, morton=morton)[0]
            for algo in algos:
                if s and algo == "chunk":
                    continue
                if s:
                    got = V.voxelize_solid_table(info, out, morton=morton, algo=algo)
                else:
                    got = V.voxelize_surface(info, out, morton=morton, algo=algo)
                assert (got.cpu().numpy().view(np.uint32) == exp).all(), (morton, s, algo)


def test_sort_triangles_torus_512(oracle, gpu):
    from cuda_voxelizer_amd import meshgen
    _check_sorted(oracle, gpu, meshgen.torus_slivers(grid=512), 512, layouts=(False, True), solid=(False,))


def test_sort_triangles_interleaved_icosphere(oracle, gpu):
    """Subdivision order scatters consecutive faces over the whole sphere."""
    from cuda_voxelizer_amd import meshgen
    v, f = meshgen.icosphere_sh(level=7)
    _check_sorted(oracle, gpu, meshgen.to_aos(v, f), 256, layouts=(False, True))


@pytest.mark.parametrize("grid,n", [(1, 1000), (16, 5000), (33, 70001), (100, 1), (2048, 3000)])
def test_sort_triangles_grids_and_counts(oracle, gpu, grid, n):
    """One brick (8 buckets, the minimum), partial last bricks, a ragged last
    wave and workgroup, a single triangle, and a grid above 1024 (32-voxel
    bricks)."""
    from cuda_voxelizer_amd import meshgen
    tris = meshgen.triangle_soup(n, seed=grid + n, mu=-4.0, sigma=1.0)
    _check_sorted(oracle, gpu, tris, grid, algos=("chunk",) if grid < 2048 else ("auto",),
                  solid=(False,) if grid == 2048 else (False, True))


def test_sort_triangles_back_to_back_and_side_stream(oracle, gpu):
    """The per-stream counts must be zero again after every call: sizes and
    bucket counts change from call to call, and one call runs on a side stream."""
    import torch
    import cuda_voxelizer_amd as V
    from cuda_voxelizer_amd import meshgen
    v, f = meshgen.icosphere_sh(level=6)
    ico = meshgen.to_aos(v, f)
    soup = meshgen.triangle_soup(120_000, seed=4, mu=-5.0, sigma=1.0)
    for tris, grid in ((ico, 512), (soup, 16), (soup, 1000), (ico, 64)):
        _check_sorted(oracle, gpu, tris, grid, solid=(False,), algos=("chunk",))
    s = torch.cuda.Stream()
    info = V.voxinfo_for_mesh(ico.reshape(-1, 3), 256, ico.shape[0])
    d = torch.from_numpy(ico.reshape(-1)).to(gpu)
    out = V.sort_triangles(info, d, stream=s)
    s.synchronize()
    got = V.voxelize_surface(info, out, algo="chunk").cpu().numpy().view(np.uint32)
    exp, _ = oracle.surface(oracle.info_from_bytes(info.as_bytes()), ico)
    assert (got == exp).all()


def test_sort_triangles_rejects_overlap_and_sorts_nothing(gpu):
    import ctypes
    import torch
    import cuda_voxelizer_amd as V
    from cuda_voxelizer_amd import _capi, meshgen
    tris = meshgen.triangle_soup(1000, seed=2)
    info = V.voxinfo_for_mesh(tris.reshape(-1, 3), 64, tris.shape[0])
    d = torch.from_numpy(tris.reshape(-1)).to(gpu)
    lib = _capi.load()
    with pytest.raises(_capi.VoxError, match="overlaps"):
        _capi.check(lib.vox_sort_triangles(ctypes.byref(info._c), d.data_ptr(), d.data_ptr() + 36 * 500, None, None))
    out = torch.empty_like(d)
    with pytest.raises(_capi.VoxError, match="overlaps"):
        _capi.check(lib.vox_sort_triangles(ctypes.byref(info._c), d.data_ptr(), out.data_ptr(), out.data_ptr(), None))
    empty = V.VoxInfo(V.AABox([0, 0, 0], [1, 1, 1]), (32, 32, 32), 0)
    assert V.sort_triangles(empty, d).numel() == 0


def test_prepare_mesh_spatial_order(oracle, gpu):
    """prepare_mesh(..., spatial_order=True): the same info, the triangles of the
    packed mesh in brick order, and the oracle's tables in both modes."""
    import cuda_voxelizer_amd as V
    from cuda_voxelizer_amd import meshgen
    v, f = meshgen.icosphere_sh(level=5, seed=9)
    info, tris = V.prepare_mesh(v, f, 128, spatial_order=True)
    assert info.as_bytes() == V.voxinfo_for_mesh(v, 128, f.shape[0]).as_bytes()
    got_tris = tris.cpu().numpy().reshape(-1, 3, 3)
    ref = meshgen.to_aos(v, f)
    key = lambda a: np.sort(a.reshape(len(a), -1).view(np.uint32).view([("", np.uint32)] * 9), axis=0)
    assert (key(got_tris) == key(ref)).all()  # the same multiset of triangles
    assert (np.diff(_brick_keys(got_tris, info)) >= 0).all()
    oi = oracle.info_from_bytes(info.as_bytes())
    assert (V.voxelize_surface(info, tris).cpu().numpy().view(np.uint32) == oracle.surface(oi, (v, f))[0]).all()
    assert (V.voxelize_solid_table(info, tris).cpu().numpy().view(np.uint32) == oracle.solid(oi, (v, f))[0]).all()


def test_sort_triangles_non_finite_vertices(gpu):
    """Vertices that are NaN or infinite still give a permutation of the input,
    bit for bit (their keys are clamped or masked into the bucket range)."""
    import torch
    import cuda_voxelizer_amd as V
    from cuda_voxelizer_amd import meshgen
    tris = meshgen.triangle_soup(5000, seed=8, mu=-4.0)
    info = V.voxinfo_for_mesh(tris.reshape(-1, 3), 256, tris.shape[0])  # box of the finite mesh
    bad = tris.copy()
    bad[::7, 0, 0] = np.nan
    bad[3::11, 1, 1] = np.inf
    bad[5::13, 2, 2] = -np.inf
    d = torch.from_numpy(bad.reshape(-1)).to(gpu)
    out, perm = V.sort_triangles(info, d, with_perm=True)
    p = perm.cpu().numpy().astype(np.int64)
    assert (np.sort(p) == np.arange(len(bad))).all()
    assert (out.cpu().numpy().reshape(-1, 3, 3).view(np.uint32) == bad[p].view(np.uint32)).all()
