"""GPU test of vox_solid_noalloc, the allocation-free solid entry point: called
through the C-ABI, its table and out-of-row count equal vox_solid's (linear
layout, marker kernels) bit for bit, on two calls in a row with the second
input smaller, including more spilled runs than its fixed scratch holds
records for; grids its scratch cannot serve are refused."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _spill_columns(grid, n_side, x_world):
    """One small triangle per column (y, z) of an n_side x n_side block, all in
    the plane x = x_world beyond the grid: each hit column's run end is past
    the row, so every triangle records a spilled run."""
    tris = []
    for y in range(n_side):
        for z in range(n_side):
            yz = np.array([[y + 0.2, z + 0.2], [y + 0.8, z + 0.3], [y + 0.4, z + 0.8]], np.float64)
            tris.append(np.concatenate([np.full((3, 1), x_world), yz], 1))
    return np.stack(tris).astype(np.float32)


def _call(capi, name, info, d_tris, table, offrow, stream):
    from cuda_voxelizer_amd._capi import VOX_LAYOUT_LINEAR, check
    if name == "vox_solid":
        check(capi.vox_solid(ctypes.byref(info._c), d_tris.data_ptr(), table.data_ptr(), VOX_LAYOUT_LINEAR, 0,
                             offrow.data_ptr(), stream))
    else:
        check(capi.vox_solid_noalloc(ctypes.byref(info._c), d_tris.data_ptr(), table.data_ptr(),
                                     offrow.data_ptr(), stream))


def _cases():
    import cuda_voxelizer_amd as V
    from cuda_voxelizer_amd import meshgen
    v, f = meshgen.icosphere_sh(level=6, seed=0x7A)
    closed = meshgen.to_aos(v, f)
    big_info = V.voxinfo_for_mesh(closed.reshape(-1, 3), 256, closed.shape[0])
    # the closed mesh, then 90,000 columns whose runs leave the row (more than
    # the 65,536 spill records: the overflow is recomputed from the triangles)
    G2 = 512
    spills = _spill_columns(G2, 300, 1.5 * G2)
    mixed = np.concatenate([closed * np.float32(G2 / 4) + np.float32(G2 / 2), spills])
    spill_info = V.VoxInfo(V.AABox(np.zeros(3, np.float32), np.full(3, G2, np.float32)), G2, mixed.shape[0])
    v2, f2 = meshgen.icosphere_jitter(level=3, seed=0x7B)
    small = meshgen.to_aos(v2, f2)
    small_info = V.voxinfo_for_mesh(small.reshape(-1, 3), 64, small.shape[0])
    return [(big_info, closed), (small_info, small), (spill_info, mixed), (small_info, small)]


def test_noalloc_matches_vox_solid(capi, gpu, oracle):
    import torch
    stream = torch.cuda.current_stream().cuda_stream
    assert capi.vox_noalloc_scratch_bytes() == 256 + 16 * 65536
    for info, tris in _cases():
        d = torch.from_numpy(np.ascontiguousarray(tris).reshape(-1)).to(gpu)
        words = capi.vox_table_bytes(info.grid, info.grid, info.grid) // 4
        ref = torch.full((words,), -1, dtype=torch.int32, device=gpu)
        got = torch.full((words,), -1, dtype=torch.int32, device=gpu)
        off_ref = torch.zeros(1, dtype=torch.int64, device=gpu)
        off_got = torch.zeros(1, dtype=torch.int64, device=gpu)
        _call(capi, "vox_solid", info, d, ref, off_ref, stream)
        _call(capi, "vox_solid_noalloc", info, d, got, off_got, stream)
        torch.cuda.synchronize()
        a = ref.cpu().numpy().view(np.uint32)
        b = got.cpu().numpy().view(np.uint32)
        assert int(np.unpackbits((a ^ b).view(np.uint8)).sum()) == 0, (info.grid, tris.shape[0])
        assert int(off_got.item()) == int(off_ref.item())
        exp, ocl, _ = oracle.solid(oracle.info_from_bytes(info.as_bytes()), tris)
        assert (b == exp).all(), (info.grid, tris.shape[0])
        assert int(off_got.item()) == ocl
    # the spill case really overflowed the record list
    assert _cases()[2][1].shape[0] > 65536


def test_noalloc_refuses_grids_it_cannot_serve(capi, gpu):
    import torch
    import cuda_voxelizer_amd as V
    from cuda_voxelizer_amd._capi import VOX_ERR_UNSUPPORTED
    tris = np.zeros((1, 3, 3), np.float32)
    d = torch.from_numpy(tris.reshape(-1)).to(gpu)
    for grid in (40, 16):
        info = V.VoxInfo(V.AABox(np.zeros(3, np.float32), np.ones(3, np.float32)), grid, 1)
        table = torch.zeros(V.table_words(grid), dtype=torch.int32, device=gpu)
        st = capi.vox_solid_noalloc(ctypes.byref(info._c), d.data_ptr(), table.data_ptr(), None,
                                    torch.cuda.current_stream().cuda_stream)
        assert st == VOX_ERR_UNSUPPORTED
