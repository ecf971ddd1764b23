"""The HIP solid kernels at the reference's thresholds: every solid algorithm,
both layouts, against the hand-derived known answers of tests/kats.py (the
|n.x| skip, the |t| on-edge test and the top-left rule of
cpu_voxelizer.cpp:196-238,265 / voxelize_solid.cu:13-92)."""
import numpy as np
import pytest

from tests import kats

pytestmark = pytest.mark.gpu

SOLID_ALGOS = ["auto", "marker", "tile", "direct", "slab", "region"]
FAMILIES = {"normal": kats.normal_threshold_cases, "edge_t": kats.edge_threshold_cases,
            "edge_ties": kats.edge_tie_cases}


def _occupancy(words: np.ndarray, morton: bool) -> np.ndarray:
    """[z, y, x] occupancy of a table, decoded independently of the oracle."""
    bits = np.unpackbits(words.astype(">u4").view(np.uint8)).astype(bool)[: kats.GRID ** 3]
    if not morton:
        return bits.reshape(kats.GRID, kats.GRID, kats.GRID)
    occ = np.zeros((kats.GRID,) * 3, bool)
    for idx in np.flatnonzero(bits):
        x = y = z = 0
        for b in range(15):  # Morton code: bit 3k is x_k, 3k+1 is y_k, 3k+2 is z_k
            x |= ((int(idx) >> (3 * b)) & 1) << b
            y |= ((int(idx) >> (3 * b + 1)) & 1) << b
            z |= ((int(idx) >> (3 * b + 2)) & 1) << b
        occ[z, y, x] = True
    return occ


@pytest.mark.parametrize("family", list(FAMILIES))
@pytest.mark.parametrize("morton", [False, True])
@pytest.mark.parametrize("algo", SOLID_ALGOS)
def test_solid_threshold_known_answers(gpu, algo, morton, family):
    import cuda_voxelizer_amd as V
    import torch
    box = V.AABox(np.array(kats.BOX_MIN, np.float32), np.array(kats.BOX_MAX, np.float32))
    info = V.VoxInfo(box, kats.GRID, 1)
    for case in FAMILIES[family]():
        print(f"case {case.name} algo={algo} morton={morton}", flush=True)
        d = torch.from_numpy(np.ascontiguousarray(case.tri.reshape(-1))).to(gpu)
        offrow = torch.zeros(1, dtype=torch.int64, device=gpu)
        t = V.voxelize_solid_table(info, d, morton=morton, algo=algo, offrow=offrow)
        torch.cuda.synchronize()
        got = _occupancy(t.cpu().numpy().view(np.uint32), morton)
        bad = np.argwhere(got != case.occupancy())
        assert bad.size == 0, (case.name, bad[:8].tolist())
        assert int(offrow.item()) == 0
