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
    algos = ["auto", "tile", "chunk", "direct", "marker"]
    for i, m in enumerate(meshes):
        morton = (m.grid & (m.grid - 1)) == 0 and i % 2 == 0
        algo = algos[i % 5]
        if (m.solid and algo == "chunk") or (not m.solid and algo == "marker"):
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


def test_graph_replays_match_the_eager_batch(oracle, gpu):
    """The graph runs its classes as concurrent branches (tile chain, marker
    chain and direct classes beside the chunk classes).  Replayed four times,
    with eager work and synchronisations of the program's own between the
    replays, every replay's tables and out-of-row counts equal those of the
    same batch launched eagerly (classes in stream order), which equal the
    oracle's."""
    import torch
    import cuda_voxelizer_amd as V
    jobs, host = _jobs(gpu, count=30, seed=0xB5)
    eager = V.Batch(jobs, use_graph=False)
    info = eager.info()
    assert info["chunk_jobs"] > 0 and info["tile_surface_jobs"] + info["tile_solid_jobs"] > 0
    assert info["direct_jobs"] > 0
    eager.launch()
    _check(oracle, jobs, host)
    ref = [j.table.clone() for j in jobs]
    ref_off = [int(j.offrow.item()) if j.offrow is not None else None for j in jobs]
    graph = V.Batch(jobs, use_graph=True)
    other = torch.zeros(1 << 20, dtype=torch.int32, device=gpu)
    for rep in range(4):
        for j in jobs:
            j.table.fill_(-1)
            if j.offrow is not None:
                j.offrow.zero_()
        other.add_(rep + 1)  # unrelated eager work between the replays
        torch.cuda.synchronize()
        print(f"graph replay {rep}", flush=True)
        graph.launch()
        torch.cuda.synchronize()
        for j, t, o in zip(jobs, ref, ref_off):
            assert torch.equal(j.table, t), (rep, j.info, j.morton, j.solid, j.algo)
            if o is not None:
                assert int(j.offrow.item()) == o, (rep, j.algo)
    assert int(other[0].item()) == 10
    graph.close()
    eager.close()


@pytest.mark.parametrize("use_graph", [True, False])
def test_empty_and_one_branch_batches(oracle, gpu, use_graph):
    """Batches whose graph has empty branches: no jobs at all, only tile jobs
    (the branch that needs no fill), only direct jobs, only chunk jobs (the
    main stream's own branch); each launches twice and matches the oracle."""
    import torch
    import cuda_voxelizer_amd as V
    from cuda_voxelizer_amd import meshgen
    empty = V.Batch([], use_graph=use_graph)
    empty.launch()
    empty.launch()
    torch.cuda.synchronize()
    assert empty.info()["jobs"] == 0
    empty.close()
    v, f = meshgen.icosphere_sh(level=3, seed=7)
    tris = meshgen.to_aos(v, f)
    for algo, solid in (("tile", False), ("tile", True), ("direct", False), ("direct", True), ("chunk", False)):
        print(f"one-branch batch algo={algo} solid={solid} graph={use_graph}", flush=True)
        jobs, host = [], []
        for g in (32, 64):
            info = V.voxinfo_for_mesh(tris.reshape(-1, 3), g, tris.shape[0])
            d = torch.from_numpy(np.ascontiguousarray(tris).reshape(-1)).to(gpu)
            t = torch.full((V.table_words(g),), -1, dtype=torch.int32, device=gpu)
            jobs.append(V.BatchJob(info, d, t, solid=solid, algo=algo))
            host.append(tris)
        b = V.Batch(jobs, use_graph=use_graph)
        b.launch()
        _check(oracle, jobs, host)
        for j in jobs:
            j.table.fill_(-1)
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


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_marker_spills_per_job(oracle, gpu, use_graph):
    """Solid jobs of the marker class whose runs leave their rows: the spill
    records carry the job index, so each spill lands in its own job's table
    (jobs of different grids, with and without spills, and offrow counts)."""
    import torch
    import cuda_voxelizer_amd as V
    from cuda_voxelizer_amd import meshgen
    jobs, host, cases = [], [], []
    for k, grid in enumerate((32, 64, 96, 32)):
        box = V.AABox(np.zeros(3, np.float32), np.full(3, grid, np.float32))
        tris = []
        for c in (grid + 1.7, 2.2 * grid, 0.2, 0.4 * grid, -0.7):
            for yz in (np.array([[0.2, 0.3], [3.1, 0.6], [0.9, 3.4]]) + k,
                       np.array([[0.2, 0.3], [grid - 0.9, 0.6], [0.9, grid - 0.6]]),
                       np.array([[grid - 2.8, grid - 2.7], [grid - 0.1, grid - 2.4], [grid - 2.1, grid - 0.1]])):
                tris.append(np.concatenate([np.full((3, 1), c), yz], 1))
        tris = np.stack(tris).astype(np.float32)
        cases.append((V.VoxInfo(box, grid, tris.shape[0]), tris))
    v, f = meshgen.icosphere_sh(level=4, seed=3)  # a closed mesh without spills in the same class
    sh = meshgen.to_aos(v, f)
    cases.insert(2, (V.voxinfo_for_mesh(sh.reshape(-1, 3), 64, sh.shape[0]), sh))
    empty = np.zeros((0, 3, 3), np.float32)  # no triangles: the table is still defined (all zero)
    cases.insert(1, (V.VoxInfo(V.AABox(np.zeros(3, np.float32), np.ones(3, np.float32)), 32, 0), empty))
    for info, tris in cases:
        d = torch.from_numpy(np.ascontiguousarray(tris).reshape(-1)).to(gpu)
        t = torch.full((V.table_words(info.grid),), -1, dtype=torch.int32, device=gpu)
        off = torch.zeros(1, dtype=torch.int64, device=gpu)
        jobs.append(V.BatchJob(info, d, t, solid=True, algo="marker", offrow=off))
        host.append(tris)
    b = V.Batch(jobs, use_graph=use_graph)
    for _ in range(2):  # the second launch reuses the spill workspace
        for j in jobs:
            j.table.fill_(-1)
            j.offrow.zero_()
        b.launch()
        _check(oracle, jobs, host)
        for j, tris in zip(jobs, host):
            _, ocl, _ = oracle.solid(oracle.info_from_bytes(j.info.as_bytes()), tris)
            assert int(j.offrow.item()) == ocl
    b.close()


def test_batch_chunk_work_shapes(oracle, gpu):
    """One chunk launch holding every work shape: small dense tables with several
    chunks per workgroup (span), coarse meshes with several workgroups per chunk
    (parts) and plain one-chunk workgroups."""
    import torch
    import cuda_voxelizer_amd as V
    from cuda_voxelizer_amd import meshgen
    fine = np.concatenate([meshgen.torus_slivers(nu=256, nv=128, grid=32, seed=1),
                           meshgen.triangle_soup(20000, seed=2, mu=-3.0, sigma=1.0)])
    coarse = meshgen.triangle_soup(700, seed=3, mu=-2.0, sigma=0.8)
    cases = [(fine, 32, True), (fine, 40, False), (coarse, 128, True), (fine, 64, False), (coarse, 96, False)]
    jobs, host = [], []
    for tris, grid, morton in cases:
        info = V.voxinfo_for_mesh(tris.reshape(-1, 3), grid, tris.shape[0])
        d = torch.from_numpy(np.ascontiguousarray(tris).reshape(-1)).to(gpu)
        t = torch.full((V.table_words(grid),), -1, dtype=torch.int32, device=gpu)
        jobs.append(V.BatchJob(info, d, t, morton=morton, algo="chunk"))
        host.append(tris)
    for use_graph in (True, False):
        b = V.Batch(jobs, use_graph=use_graph)
        b.launch()
        _check(oracle, jobs, host)
        for j in jobs:
            j.table.fill_(-1)
        b.launch()
        _check(oracle, jobs, host)
        b.close()


def test_c4_batch_256_meshes_one_graph_launch(oracle, gpu, tmp_path):
    """The batched workload itself: 256 meshes (seed 0xC4; triangle counts
    log-uniform in [500, 500k], ~16M triangles in all; grids {32, 64, 128, 256};
    icospheres and tori solid, soups surface), with the per-class algorithms
    the autotuner picks, replayed as one graph launch; every table equals the
    oracle's, on the first launch and on a relaunch over dirty tables."""
    import torch
    import cuda_voxelizer_amd as V
    from cuda_voxelizer_amd import meshgen
    from cuda_voxelizer_amd.autotune import Tuner
    meshes = meshgen.batch_meshes(256, seed=0xC4)
    jobs, host = [], []
    for m in meshes:
        info = V.voxinfo_for_mesh(m.tris.reshape(-1, 3), m.grid, m.tris.shape[0])
        d = torch.from_numpy(m.tris.reshape(-1)).to(gpu)
        t = torch.full((V.table_words(m.grid),), -1, dtype=torch.int32, device=gpu)
        jobs.append(V.BatchJob(info, d, t, solid=m.solid))
        host.append(m.tris)
    Tuner(path=str(tmp_path / "autotune.json")).tune_jobs(jobs)
    b = V.Batch(jobs, use_graph=True)
    assert b.info()["graph"] == 1 and b.info()["jobs"] == 256
    b.launch()
    _check(oracle, jobs, host)
    for j in jobs:
        j.table.fill_(-1)
    b.launch()
    _check(oracle, jobs, host)
    b.close()


def test_tuning_knobs_same_tables(oracle, gpu):
    """Every knob setting the autotuner may pick gives the oracle's table, in a
    batch (both chunk classes, marker and tile jobs side by side) and in plain
    calls: chunk waves 5/6 and queue thresholds 16..4096 voxels, marker
    thresholds 1..1024 columns, tile sizes half/twice."""
    import torch
    import cuda_voxelizer_amd as V
    from cuda_voxelizer_amd import _capi, meshgen
    soup = meshgen.triangle_soup(6000, seed=0x71, mu=-3.0, sigma=1.0)
    torus = meshgen.torus_slivers(nu=128, nv=64, grid=128, seed=0x72)
    v, f = meshgen.icosphere_sh(level=4, seed=0x73)
    closed = meshgen.to_aos(v, f)
    coarse = meshgen.to_aos(*meshgen.icosphere_jitter(level=1, seed=0x74))
    cases = [(soup, 128, False, "chunk", _capi.tune_waves(1) | _capi.tune_big(4)),
             (soup, 128, False, "chunk", _capi.tune_waves(2) | _capi.tune_big(12)),
             (torus, 256, False, "chunk", _capi.tune_waves(2)),
             (torus, 64, True, "marker", _capi.tune_big(0)),
             (closed, 128, True, "marker", _capi.tune_big(10)),
             (coarse, 128, True, "tile", _capi.tune_tile(1)),
             (coarse, 256, False, "tile", _capi.tune_tile(3)),
             (soup, 96, False, "tile", _capi.tune_tile(3))]
    jobs, host = [], []
    for tris, grid, solid, algo, tune in cases:
        info = V.voxinfo_for_mesh(tris.reshape(-1, 3), grid, tris.shape[0])
        d = torch.from_numpy(np.ascontiguousarray(tris).reshape(-1)).to(gpu)
        t = torch.full((V.table_words(grid),), -1, dtype=torch.int32, device=gpu)
        jobs.append(V.BatchJob(info, d, t, solid=solid, algo=algo, tune=tune))
        host.append(tris)
    b = V.Batch(jobs)
    b.launch()
    _check(oracle, jobs, host)
    b.close()
    for j, tris in zip(jobs, host):  # the same settings as plain calls
        j.table.fill_(-1)
        if j.solid:
            V.voxelize_solid_table(j.info, j.tris, table=j.table, algo=j.algo, tune=j.tune)
        else:
            V.voxelize_surface(j.info, j.tris, table=j.table, algo=j.algo, tune=j.tune)
    _check(oracle, jobs, host)
