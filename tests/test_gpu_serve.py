"""GPU tests of the continuous-batching server: every result is bit-identical
to the same request run alone through the plain path (and to the oracle),
whatever micro-batch it ran in; admission respects the HBM budget."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _requests(n=24, seed=7, grids=(32, 64, 128)):
    from cuda_voxelizer_amd import meshgen
    from cuda_voxelizer_amd.serve import Request
    pool = meshgen.batch_meshes(8, seed=seed, n_min=200, n_max=20000, grids=grids)
    rng = np.random.default_rng(seed)
    reqs, t = [], 0.0
    for i in range(n):
        m = pool[int(rng.integers(len(pool)))]
        g = int(rng.choice(grids))
        solid = bool(m.solid and rng.random() < 0.5)
        t += float(rng.exponential(0.002))
        reqs.append(Request(i, id(m), m.tris, g, solid, morton=bool(rng.random() < 0.5), arrival=t))
    return reqs


def _run_server(server, reqs, clock):
    t0 = clock()
    pending = sorted(reqs, key=lambda r: r.arrival)
    out = []
    while pending or server.pending():
        now = clock() - t0
        while pending and pending[0].arrival <= now:
            server.submit(pending.pop(0))
        out += server.step()
    return out


def test_results_same_alone_and_in_the_mix(oracle, gpu):
    import time
    from cuda_voxelizer_amd.serve import VoxServer, serve_plain
    reqs = _requests()
    server = VoxServer(streams=2, max_batch=8)
    mixed = {r.rid: r for r in _run_server(server, reqs, time.perf_counter)}
    assert len(mixed) == len(reqs)
    assert server.stats["batches"] < len(reqs)  # requests really were batched together
    assert max(r.batch_size for r in mixed.values()) > 1
    alone = {r.rid: r for r in serve_plain(reqs, realtime=False)}
    for req in reqs:
        a = mixed[req.rid].table.cpu().numpy().view(np.uint32)
        b = alone[req.rid].table.cpu().numpy().view(np.uint32)
        assert (a == b).all(), req.rid
    for req in reqs[:6]:
        oi = oracle.info_from_bytes(mixed[req.rid].info.as_bytes())
        if req.solid:
            exp, _, _ = oracle.solid(oi, req.tris, morton=req.morton)
        else:
            exp, _ = oracle.surface(oi, req.tris, morton=req.morton)
        assert (mixed[req.rid].table.cpu().numpy().view(np.uint32) == exp).all(), req.rid
    server.close()


def test_budget_admission_and_rejection(gpu):
    import time
    import cuda_voxelizer_amd as V
    from cuda_voxelizer_amd.serve import Request, VoxServer
    reqs = _requests(n=12, seed=9, grids=(128,))
    table = V.table_bytes(128)
    biggest_mesh = max(r.tris.size * 4 for r in reqs)
    budget = 2 * table + 2 * biggest_mesh + 1
    server = VoxServer(budget_bytes=budget, streams=2, max_batch=4)
    peak = 0
    t0 = time.perf_counter()
    pending = list(reqs)
    done = []
    while pending or server.pending():
        while pending:
            server.submit(pending.pop(0))
        for r in server.step():
            done.append(r)
            server.release(r)
        peak = max(peak, server.used)
        assert time.perf_counter() - t0 < 120
    assert len(done) == len(reqs)
    assert peak <= budget
    huge = Request(99, "huge", reqs[0].tris, 4096, False)
    server.submit(huge)
    server.step()
    assert 99 in server.rejected
    server.close()


def _check_against(oracle, mixed, reqs, alone=None):
    for req in reqs:
        got = mixed[req.rid].table.cpu().numpy().view(np.uint32)
        if alone is not None:
            assert (got == alone[req.rid].table.cpu().numpy().view(np.uint32)).all(), req.rid
        oi = oracle.info_from_bytes(mixed[req.rid].info.as_bytes())
        if req.solid:
            exp, _, _ = oracle.solid(oi, req.tris, morton=req.morton)
        else:
            exp, _ = oracle.surface(oi, req.tris, morton=req.morton)
        assert (got == exp).all(), (req.rid, req.grid, req.solid, req.morton)


def test_large_mesh_uploaded_mid_run(oracle, gpu):
    """A ~500k-triangle mesh arrives while both runners are busy; its upload
    (18 MB) is still in flight when its next requests are launched on the other
    runner.  Every launch waits for the mesh's ready event, so each table is
    the oracle's."""
    import os
    import time
    from cuda_voxelizer_amd import meshgen
    from cuda_voxelizer_amd.serve import Request, VoxServer
    oracle.set_threads(min(16, os.cpu_count() or 1))
    v, f = meshgen.icosphere_jitter(level=7, seed=0x5E)  # 327,680 triangles
    big = meshgen.to_aos(v, f)
    big = np.concatenate([big, meshgen.triangle_soup(170_000, seed=0x5F)])  # ~500k triangles, 17.9 MB
    small = meshgen.batch_meshes(4, seed=0x60, n_min=2000, n_max=20000, grids=(128,))
    reqs = []
    for i, m in enumerate(small):  # keep both runners busy first
        reqs.append(Request(i, f"s{i}", m.tris, 256, m.solid))
    for k in range(6):  # then the large mesh, first seen now, on alternating runners
        reqs.append(Request(10 + k, "big", big, (128, 256)[k % 2], k % 3 == 0, morton=k % 2 == 1))
    server = VoxServer(streams=2, max_batch=1)
    for r in reqs:
        server.submit(r)
    out = {}
    t0 = time.perf_counter()
    while server.pending():
        for res in server.step():
            out[res.rid] = res
        assert time.perf_counter() - t0 < 120
    assert server.stats["mesh_uploads"] == len(small) + 1
    _check_against(oracle, out, reqs)
    server.close()


def _c5_requests(rate, n, pool, rng):
    from cuda_voxelizer_amd.serve import Request
    reqs, t = [], 0.0
    for i in range(n):
        k = int(rng.integers(len(pool)))
        m = pool[k]
        g = int(rng.choice((64, 128, 256, 512)))
        solid = bool(m.solid and rng.random() < 0.5)  # soups stay surface
        t += float(rng.exponential(1.0 / rate))
        reqs.append(Request(i, ("c5", k), m.tris, g, solid, arrival=t))
    return reqs


@pytest.mark.parametrize("rate", [150.0, 600.0, 2400.0])
def test_c5_mix_same_alone_and_in_the_mix(oracle, gpu, rate):
    """The serving workload: requests drawn from the batched-mesh distribution
    (log-uniform 500..500k triangles, icospheres / tori / soups), grids
    {64,128,256,512}, surface or solid 50/50, Poisson arrivals at three rates,
    64 GiB budget; meshes are uploaded during the run when first requested.
    Each table is the same alone (plain per-request call) and in the mix, and
    equals the oracle."""
    import os
    import time
    from cuda_voxelizer_amd import meshgen
    from cuda_voxelizer_amd.serve import VoxServer, serve_plain
    oracle.set_threads(min(16, os.cpu_count() or 1))
    pool = meshgen.batch_meshes(12, seed=0xC5)
    rng = np.random.default_rng(int(rate))
    reqs = _c5_requests(rate, 40, pool, rng)
    server = VoxServer(budget_bytes=64 << 30, streams=2, max_batch=64)
    mixed = {r.rid: r for r in _run_server(server, reqs, time.perf_counter)}
    assert len(mixed) == len(reqs)
    assert server.stats["mesh_uploads"] >= 1
    alone = {r.rid: r for r in serve_plain(reqs, realtime=False)}
    _check_against(oracle, mixed, reqs, alone)
    server.close()
