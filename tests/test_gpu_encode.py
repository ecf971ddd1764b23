"""Device encoders (vox_encode) against the oracle's restatement of the
reference writers (util_io.cpp), byte for byte."""
import numpy as np
import pytest

from tests.test_writers_oracle import pack, random_vox

pytestmark = pytest.mark.gpu

FORMATS = [("binvox", "binvox_body"), ("obj_points", "obj_points"), ("obj", "obj_cubes"), ("vox", "vox_file")]


@pytest.mark.parametrize("G,p", [(1, 1.0), (5, 0.3), (8, 0.02), (24, 0.1), (32, 0.01), (64, 0.003), (96, 0.5),
                                 (128, 0.0005), (256, 0.00002)])
@pytest.mark.parametrize("fmt,ofn", FORMATS)
def test_encoder_matches_oracle_linear(oracle, gpu, G, p, fmt, ofn):
    from cuda_voxelizer_amd import io
    vox = random_vox(G, G + 1, p)
    if fmt != "binvox":
        vox[: max(1, G // 4)] = False  # keep text outputs small
    table = pack(vox)
    print(f"encode {fmt} G={G}", flush=True)
    assert io.encode(table, G, fmt) == getattr(oracle, ofn)(table, G)


@pytest.mark.parametrize("G", [4, 16, 64])
@pytest.mark.parametrize("fmt,ofn", FORMATS)
def test_encoder_matches_oracle_morton(oracle, gpu, G, fmt, ofn):
    from cuda_voxelizer_amd import io
    vox = random_vox(G, 11 * G, 0.02)
    vox[: G // 4] = False
    table = pack(vox, morton=True)
    print(f"encode morton {fmt} G={G}", flush=True)
    assert io.encode(table, G, fmt, morton=True) == getattr(oracle, ofn)(table, G, morton=True)


def test_encoder_on_voxelized_bunny(oracle, gpu):
    import os
    import cuda_voxelizer_amd as V
    from cuda_voxelizer_amd import io
    from tests.conftest import GOLDEN
    g = np.load(os.path.join(GOLDEN, "bunny.npz"))
    v, f = g["vertices"], g["faces"]
    info = V.voxinfo_for_mesh(v, 64, f.shape[0])
    import torch
    from cuda_voxelizer_amd import meshgen
    d = torch.from_numpy(meshgen.to_aos(v, f).reshape(-1)).to(gpu)
    with io.Encoder() as enc:  # one encoder, reused across formats and tables
        for solid in (False, True):
            t = (V.voxelize_solid_table if solid else V.voxelize_surface)(info, d)
            host = t.cpu().numpy().view(np.uint32)
            for fmt, ofn in FORMATS:
                assert enc.encode(t, 64, False, io.FORMATS[fmt]) == getattr(oracle, ofn)(host, 64)


def test_binvox_header_format(gpu):
    import cuda_voxelizer_amd as V
    from cuda_voxelizer_amd import io
    bbox = V.AABox(np.array([-1.25, 0.5, 3.0e-7], np.float32), np.array([2.0, 1.0, 1.0], np.float32))
    info = V.VoxInfo(bbox, 64, 10)
    h = io.binvox_header(info).decode()
    mn = [float(np.float32(x)) for x in (-1.25, 0.5, 3.0e-7)]
    scale = float(np.float32(2.0) - np.float32(-1.25))
    assert h == "#binvox 1\ndim 64 64 64\ntranslate %g %g %g\nscale %g\ndata\n" % (*mn, scale)


def test_encoder_rejects_bad_arguments(gpu):
    import torch
    from cuda_voxelizer_amd import io
    from cuda_voxelizer_amd._capi import VoxError
    t = torch.zeros(8, dtype=torch.int32, device=gpu)
    with pytest.raises(VoxError):
        io.encode(t, 4096, "binvox")
    with pytest.raises(VoxError):
        io.encode(t, 6, "binvox", morton=True)


def test_binvox_encoder_reuse_grows(oracle, gpu):
    """One encoder over tables whose pair counts rise and fall: the pairs are
    emitted into the buffer a previous call left before their number is known,
    and emitted again into a larger one when they do not fit."""
    from cuda_voxelizer_amd import io
    with io.Encoder() as enc:
        for G, p in ((64, 0.001), (64, 0.5), (128, 0.4), (32, 0.0), (96, 0.02), (128, 0.45)):
            table = pack(random_vox(G, G + 7, p))
            print(f"encode binvox reuse G={G} p={p}", flush=True)
            assert enc.encode(table, G, False, io.FORMATS["binvox"]) == oracle.binvox_body(table, G)
