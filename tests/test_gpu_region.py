"""GPU parity of the region algorithm's own paths, against the oracle.

The region algorithm stages (word, mask) pairs per output region and builds
each region once; pairs that find no room (staging buffer or a region's
descriptor list full, LDS hash crowded) go to a scratch table that the build
pass merges and clears.  The capacity bits (VOX_REGION_*) shrink the
buffers so that every fallback route runs; any setting must give the oracle's
table, and the scratch table must be left clean for the next call.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(oracle, gpu, tris, grid, morton, solid, tune=0, table=None, accumulate=False):
    import torch
    import cuda_voxelizer_amd as V
    tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 3, 3)
    info = V.voxinfo_for_mesh(tris.reshape(-1, 3), grid, tris.shape[0])
    d = torch.from_numpy(tris.reshape(-1)).to(gpu)
    oinfo = oracle.info_from_bytes(info.as_bytes())
    base = None if table is None else table.copy()
    t = None if table is None else torch.from_numpy(table.view(np.int32).copy()).to(gpu)
    if solid:
        off = torch.zeros(1, dtype=torch.int64, device=gpu)
        t = V.voxelize_solid_table(info, d, morton=morton, table=t, accumulate=accumulate, algo="region",
                                   offrow=off, tune=tune)
        exp, ocl, _ = oracle.solid(oinfo, tris, morton=morton, table=base)
        torch.cuda.synchronize()
        assert int(off.item()) == ocl
    else:
        t = V.voxelize_surface(info, d, morton=morton, table=t, accumulate=accumulate, algo="region", tune=tune)
        exp, _ = oracle.surface(oinfo, tris, morton=morton, table=base)
    got = t.cpu().numpy().view(np.uint32)
    return int(np.unpackbits((got ^ exp).view(np.uint8)).sum()), int(np.unpackbits(exp.view(np.uint8)).sum())


def _knobs():
    from cuda_voxelizer_amd import _capi
    return {"builtin": 0,
            "stage16": _capi.region_stage(4),            # the staging buffer holds 16 pairs
            "desc1": _capi.region_desc(0),               # one descriptor per region
            "words1k": _capi.region_words(10),           # 1024-word regions
            "all": _capi.region_stage(10) | _capi.region_desc(1) | _capi.region_words(10)}


@pytest.mark.parametrize("knob", ["builtin", "stage16", "desc1", "words1k", "all"])
@pytest.mark.parametrize("solid", [False, True])
@pytest.mark.parametrize("morton", [False, True])
def test_region_capacity_fallbacks(oracle, gpu, knob, solid, morton):
    from cuda_voxelizer_amd import meshgen
    v, f = meshgen.icosphere_jitter(level=5, amp=0.05, seed=21)
    bad, total = _run(oracle, gpu, meshgen.to_aos(v, f), 256, morton, solid, tune=_knobs()[knob])
    assert bad == 0 and total > 10000


@pytest.mark.parametrize("grid,level", [(512, 1), (256, 2), (128, 0)])
@pytest.mark.parametrize("solid", [False, True])
def test_region_coarse_meshes(oracle, gpu, grid, level, solid):
    """Large triangles: many words per workgroup, so the LDS hash is crowded and
    the scratch fallback carries part of the table."""
    from cuda_voxelizer_amd import meshgen
    v, f = meshgen.icosphere_sh(level=level, seed=grid)
    for morton in (False, True):
        bad, total = _run(oracle, gpu, meshgen.to_aos(v, f), grid, morton, solid)
        assert bad == 0 and total > 0, morton


def test_region_workspace_reuse(oracle, gpu):
    """Calls of different grids, layouts and modes on one stream share the
    workspace; a call that leaves pairs in the scratch table (tiny capacities)
    is followed by calls that must not see them; accumulation in between."""
    from cuda_voxelizer_amd import _capi, meshgen
    v, f = meshgen.icosphere_sh(level=5, seed=7)
    tris = meshgen.to_aos(v, f)
    tiny = _capi.region_stage(3) | _capi.region_desc(0)
    rng = np.random.default_rng(5)
    for grid, morton, solid, tune, acc in [(256, False, False, tiny, False), (128, True, True, 0, False),
                                           (512, True, False, tiny, False), (256, False, True, tiny, True),
                                           (64, False, False, 0, True), (256, True, True, 0, False),
                                           (256, False, False, 0, False), (128, False, True, 0, False)]:
        table = None
        if acc:
            import cuda_voxelizer_amd as V
            table = rng.integers(0, 2**32, V.table_words(grid), dtype=np.uint64).astype(np.uint32)
        bad, total = _run(oracle, gpu, tris, grid, morton, solid, tune=tune, table=table, accumulate=acc)
        assert bad == 0 and total > 0, (grid, morton, solid, tune, acc)
