"""GPU parity of the batched mode: every job of a mixed batch (uneven sizes,
grids, layouts, surface and solid, every algorithm) is bit-identical to the
oracle, through graph and eager launches, on repeated launches and after the
triangle contents change."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mismatch(a, b):
    return int(np.unpackbits((a ^ b).view(np.uint8)).sum())


def _jobs(gpu, count=24, seed=0xB1, n_max=20000):
    import torch
    import cuda_voxelizer_amd as V
    from cuda_voxelizer_amd import meshgen
    meshes = meshgen.batch_meshes(count, seed=seed, n_min=50, n_max=n_max, grids=(16, 32, 33, 64, 96, 128))
    jobs, host = [], []
    algos = ["auto", "tile", "chunk", "direct"]
    for i, m in enumerate(meshes):
        morton = (m.grid & (m.grid - 1)) == 0 and i % 2 == 0
        algo = algos[i % 4]
        if m.solid and algo == "chunk":
            algo = "auto"
        info = V.voxinfo_for_mesh(m.tris.reshape(-1, 3), m.grid, m.tris.shape[0])
        d = torch.from_numpy(m.tris.reshape(-1)).to(gpu)
        t = torch.full((V.table_words(m.grid),), -1, dtype=torch.int32, device=gpu)
        off = torch.zeros(1, dtype=torch.int64, device=gpu) if m.solid else None
        jobs.append(V.BatchJob(info, d, t, morton=morton, solid=m.solid, algo=algo, offrow=off))
        host.append(m.tris)
    return jobs, host


def _check(oracle, jobs, host):
    import torch
    torch.cuda.synchronize()
    for j, tris in zip(jobs, host):
        oi = oracle.info_from_bytes(j.info.as_bytes())
        if j.solid:
            exp, _, _ = oracle.solid(oi, tris, morton=j.morton)
        else:
            exp, _ = oracle.surface(oi, tris, morton=j.morton)
        got = j.table.cpu().numpy().view(np.uint32)
        assert _mismatch(got, exp) == 0, (j.info, j.morton, j.solid, j.algo)


@pytest.fixture(scope="module", autouse=True)
def _threads(oracle):
    oracle.set_threads(min(16, os.cpu_count() or 1))


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_matches_oracle(oracle, gpu, use_graph):
    import cuda_voxelizer_amd as V
    jobs, host = _jobs(gpu)
    b = V.Batch(jobs, use_graph=use_graph)
    info = b.info()
    assert info["jobs"] == len(jobs) and info["graph"] == (1 if use_graph else 0)
    assert info["chunk_jobs"] > 0 and info["tile_surface_jobs"] > 0 and info["tile_solid_jobs"] > 0
    b.launch()
    _check(oracle, jobs, host)
    # second launch over dirty tables: the batch redefines every table
    for j in jobs:
        j.table.fill_(-1)
    b.launch()
    b.launch()
    _check(oracle, jobs, host)
    b.close()


def test_batch_sees_new_triangle_contents(oracle, gpu):
    import torch
    import cuda_voxelizer_amd as V
    jobs, host = _jobs(gpu, count=12, seed=0xB2)
    b = V.Batch(jobs)
    b.launch()
    # shrink every mesh towards its centroid (still inside its box) and replay
    new_host = []
    for j, tris in zip(jobs, host):
        c = tris.reshape(-1, 3).mean(0)
        t2 = (c + (tris - c) * np.float32(0.7)).astype(np.float32)
        j.tris.copy_(torch.from_numpy(t2.reshape(-1)))
        new_host.append(t2)
    b.launch()
    _check(oracle, jobs, new_host)
    b.close()
