"""GPU parity of the write-once slab algorithm (linear layout): tile shapes of
every grid size it plans, triangles spanning more tiles than a list entry
holds (the finish pass), solid runs leaving their row beyond the record list
(recomputed), deep columns, accumulation, and consecutive calls on one
stream whose workspace must come back clean."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mismatch(a, b):
    return int(np.unpackbits((a ^ b).view(np.uint8)).sum())


@pytest.fixture(scope="module", autouse=True)
def _threads(oracle):
    oracle.set_threads(min(16, os.cpu_count() or 1))


def _run(oracle, gpu, tris, info, solid, table=None, accumulate=False):
    import torch
    import cuda_voxelizer_amd as V
    d = torch.from_numpy(np.ascontiguousarray(tris, np.float32).reshape(-1)).to(gpu)
    off = torch.zeros(1, dtype=torch.int64, device=gpu)
    if solid:
        t = V.voxelize_solid_table(info, d, table=table, accumulate=accumulate, algo="slab", offrow=off)
        exp, ocl, _ = oracle.solid(oracle.info_from_bytes(info.as_bytes()), tris)
    else:
        t = V.voxelize_surface(info, d, table=table, accumulate=accumulate, algo="slab")
        exp, ocl = oracle.surface(oracle.info_from_bytes(info.as_bytes()), tris)[0], 0
    torch.cuda.synchronize()
    return t.cpu().numpy().view(np.uint32), exp, int(off.item()), ocl


@pytest.mark.parametrize("grid", [32, 64, 96, 160, 256, 512, 2048])
def test_slab_grids_surface_and_solid(oracle, gpu, grid):
    import cuda_voxelizer_amd as V
    from cuda_voxelizer_amd import meshgen
    v, f = meshgen.icosphere_sh(level=5 if grid < 2048 else 6, seed=grid)
    tris = meshgen.to_aos(v, f)
    info = V.voxinfo_for_mesh(tris.reshape(-1, 3), grid, tris.shape[0])
    print(f"slab grid {grid}", flush=True)
    for solid in (False, True):
        got, exp, off, ocl = _run(oracle, gpu, tris, info, solid)
        assert _mismatch(got, exp) == 0, solid
        assert off == ocl


def test_slab_wide_triangles(oracle, gpu):
    """Large soup triangles cover many slabs (the finish pass applies them with
    global atomics after the slabs are written), mixed with small ones; and a
    coarse closed mesh for the solid finish pass."""
    import cuda_voxelizer_amd as V
    from cuda_voxelizer_amd import meshgen
    big = meshgen.triangle_soup(400, seed=11, mu=-1.5, sigma=0.6)
    tiny = meshgen.triangle_soup(30000, seed=12, mu=-4.0, sigma=0.5)
    tris = np.concatenate([tiny[:15000], big, tiny[15000:]])
    for grid in (256, 1024):
        info = V.voxinfo_for_mesh(tris.reshape(-1, 3), grid, tris.shape[0])
        got, exp, _, _ = _run(oracle, gpu, tris, info, False)
        assert _mismatch(got, exp) == 0, grid
    v, f = meshgen.icosphere_sh(level=2, seed=13)  # 320 triangles: most span several slabs
    closed = meshgen.to_aos(v, f)
    for grid in (128, 1024):
        info = V.voxinfo_for_mesh(closed.reshape(-1, 3), grid, closed.shape[0])
        got, exp, off, ocl = _run(oracle, gpu, closed, info, True)
        assert _mismatch(got, exp) == 0, grid
        assert off == ocl


def _spill_columns(n_side, x_world, y0=0.0):
    tris = []
    for y in range(n_side):
        for z in range(n_side):
            yz = np.array([[y + y0 + 0.2, z + 0.2], [y + y0 + 0.8, z + 0.3], [y + y0 + 0.4, z + 0.8]], np.float64)
            tris.append(np.concatenate([np.full((3, 1), x_world), yz], 1))
    return np.stack(tris).astype(np.float32)


@pytest.mark.parametrize("n_side", [40, 300])
def test_slab_solid_spills(oracle, gpu, n_side):
    """Runs whose end lies past the row (the plane x = 1.5 G beyond the grid):
    1,600 of them (spill records) and 90,000 (more than the 65,536 records:
    recomputed from the binned triangles), with a closed mesh in the same
    call, and runs with a negative end (x < 0: empty)."""
    import cuda_voxelizer_amd as V
    from cuda_voxelizer_amd import meshgen
    G = 512
    v, f = meshgen.icosphere_sh(level=5, seed=21)
    closed = meshgen.to_aos(v, f) * np.float32(G / 4) + np.float32(G / 2)
    tris = np.concatenate([closed, _spill_columns(n_side, 1.5 * G), _spill_columns(30, -7.0, y0=400.0)])
    info = V.VoxInfo(V.AABox(np.zeros(3, np.float32), np.full(3, G, np.float32)), G, tris.shape[0])
    got, exp, off, ocl = _run(oracle, gpu, tris, info, True)
    assert _mismatch(got, exp) == 0
    assert off == ocl and off >= n_side * n_side


def test_slab_accumulate_and_back_to_back(oracle, gpu):
    """OR / XOR into a table that holds random bits, then plain calls of
    other grids and modes on the same stream (the workspace's counters and
    counts must be left zero by each call)."""
    import torch
    import cuda_voxelizer_amd as V
    from cuda_voxelizer_amd import meshgen
    v, f = meshgen.icosphere_sh(level=4, seed=31)
    tris = meshgen.to_aos(v, f)
    for grid, solid, acc in [(128, False, True), (128, True, True), (64, True, False), (256, False, False),
                             (96, True, True), (1024, True, False), (32, False, False), (1024, False, False)]:
        info = V.voxinfo_for_mesh(tris.reshape(-1, 3), grid, tris.shape[0])
        base = np.random.default_rng(grid).integers(0, 2**32, V.table_words(grid), dtype=np.uint64).astype(np.uint32)
        t = torch.from_numpy(base.view(np.int32).copy()).to(gpu)
        got, exp, off, ocl = _run(oracle, gpu, tris, info, solid, table=t, accumulate=acc)
        if acc:
            exp = (exp ^ base) if solid else (exp | base)
        assert _mismatch(got, exp) == 0, (grid, solid, acc)


def test_slab_deep_columns(oracle, gpu):
    """Forty nested shells: up to 80 markers per column, many in one slab."""
    import cuda_voxelizer_amd as V
    from cuda_voxelizer_amd import meshgen
    v, f = meshgen.icosphere_sh(level=3, seed=41)
    tris = np.concatenate([meshgen.to_aos(v * np.float32(0.2 + 0.02 * k), f) for k in range(40)])
    info = V.voxinfo_for_mesh(tris.reshape(-1, 3), 96, tris.shape[0])
    got, exp, off, ocl = _run(oracle, gpu, tris, info, True)
    assert _mismatch(got, exp) == 0 and off == ocl
