"""CPU tests of the autotuner's bookkeeping (timing injected)."""
from types import SimpleNamespace


def _job(grid, solid, n, morton=False):
    return SimpleNamespace(info=SimpleNamespace(grid=grid, n_triangles=n), solid=solid, morton=morton,
                           algo="auto")


def test_choose_times_once_per_class_and_persists(tmp_path):
    from cuda_voxelizer_amd.autotune import Tuner, class_key
    calls = []

    def timer(job, algo):
        calls.append((job.info.grid, algo))
        return {"chunk": 3.0, "tile": 1.0, "direct": 2.0}[algo]

    path = str(tmp_path / "tune.json")
    t = Tuner(path, device_name="fake", timer=timer)
    jobs = [_job(64, False, 1000), _job(64, False, 1500), _job(64, True, 1000), _job(128, False, 1000)]
    t.tune_jobs(jobs)
    assert [j.algo for j in jobs] == ["tile", "tile", "tile", "tile"]
    # 64/surface/n2^9..10 are different buckets? 1000 -> 2^9, 1500 -> 2^10
    assert class_key(64, False, False, 1000) != class_key(64, False, False, 1500)
    n_first = len(calls)
    assert all(a in ("chunk", "tile", "direct") for _, a in calls)
    t2 = Tuner(path, device_name="fake", timer=timer)
    t2.tune_jobs([_job(64, False, 1000)])
    assert len(calls) == n_first  # served from the cache file
    t3 = Tuner(path, device_name="other", timer=timer)
    t3.tune_jobs([_job(64, False, 1000)])
    assert len(calls) > n_first  # a different device tunes afresh


def test_solid_jobs_never_get_chunk(tmp_path):
    from cuda_voxelizer_amd.autotune import Tuner
    t = Tuner(str(tmp_path / "t.json"), device_name="fake",
              timer=lambda job, algo: {"chunk": 0.1, "tile": 1.0, "direct": 2.0}[algo])
    j = _job(32, True, 100)
    assert t.choose(j) == "tile"
