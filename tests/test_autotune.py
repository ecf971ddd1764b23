"""CPU tests of the autotuner's bookkeeping (timing injected)."""
from types import SimpleNamespace


def _job(grid, solid, n, morton=False):
    return SimpleNamespace(info=SimpleNamespace(grid=grid, n_triangles=n), solid=solid, morton=morton,
                           algo="auto")


def test_choose_times_once_per_class_and_persists(tmp_path):
    from cuda_voxelizer_amd.autotune import Tuner, class_key
    calls = []

    def timer(job, algo, tune=0):
        calls.append((job.info.grid, algo))
        return {"chunk": 3.0, "marker": 4.0, "tile": 1.0, "direct": 2.0}[algo]

    path = str(tmp_path / "tune.json")
    t = Tuner(path, device_name="fake", timer=timer)
    jobs = [_job(64, False, 1000), _job(64, False, 1500), _job(64, True, 1000), _job(128, False, 1000)]
    t.tune_jobs(jobs)
    assert [j.algo for j in jobs] == ["tile", "tile", "tile", "tile"]
    # 64/surface/n2^9..10 are different buckets? 1000 -> 2^9, 1500 -> 2^10
    assert class_key(64, False, False, 1000) != class_key(64, False, False, 1500)
    n_first = len(calls)
    assert all(a in ("chunk", "marker", "tile", "direct") for _, a in calls)
    t2 = Tuner(path, device_name="fake", timer=timer)
    t2.tune_jobs([_job(64, False, 1000)])
    assert len(calls) == n_first  # served from the cache file
    t3 = Tuner(path, device_name="other", timer=timer)
    t3.tune_jobs([_job(64, False, 1000)])
    assert len(calls) > n_first  # a different device tunes afresh


def test_solid_jobs_never_get_chunk(tmp_path):
    from cuda_voxelizer_amd.autotune import Tuner
    t = Tuner(str(tmp_path / "t.json"), device_name="fake",
              timer=lambda job, algo, tune=0: {"chunk": 0.1, "marker": 3.0, "tile": 1.0, "direct": 2.0}[algo])
    j = _job(32, True, 100)
    assert t.choose(j) == "tile"


def test_copies_fill_the_device_within_the_table_budget(tmp_path):
    import numpy as np
    from cuda_voxelizer_amd.autotune import Tuner
    t = Tuner(str(tmp_path / "t.json"), device_name="fake", timer=lambda job, algo, tune=0: 1.0,
              fill_triangles=2_000_000, max_table_bytes=64 << 20)

    def job(grid, n):
        words = (grid ** 3 + 31) // 32
        return SimpleNamespace(info=SimpleNamespace(grid=grid, n_triangles=n),
                               table=SimpleNamespace(numel=lambda: words, element_size=lambda: 4))

    assert t.copies(job(64, 500)) == 64        # tiny mesh: the most copies
    assert t.copies(job(64, 100_000)) == 20    # ceil(2M / 100k)
    assert t.copies(job(64, 5_000_000)) == 4   # large mesh: at least 4
    assert t.copies(job(256, 500)) == 32       # 2 MiB tables: 64 MiB budget
    assert t.copies(job(1024, 500)) == 1       # one 128 MiB table exceeds the budget: still one copy
    assert np.all([t.copies(job(g, n)) >= 1 for g in (1, 32, 512) for n in (0, 1, 10 ** 7)])


def test_pick_rules_out_a_long_critical_path(tmp_path):
    from cuda_voxelizer_amd.autotune import Tuner
    t = Tuner(str(tmp_path / "t.json"), device_name="fake", timer=lambda job, algo, tune=0: 1.0)
    # direct: cheapest per copy, but one copy alone runs 1.3 ms (a straggler thread) -> ruled out
    times = {"tile": (60e-6, 30e-6), "direct": (1.3e-3, 20e-6)}
    assert t.pick(times) == "tile"
    # within the slack the per-copy time decides
    times = {"chunk": (80e-6, 10e-6), "tile": (50e-6, 40e-6), "direct": (300e-6, 5e-6)}
    assert t.pick(times) == "chunk"
    assert t.pick({"tile": 2.0, "direct": 1.0}) == "direct"  # plain numbers: both times


def test_direct_is_a_candidate_for_fine_meshes_only(tmp_path):
    from cuda_voxelizer_amd.autotune import Tuner
    t = Tuner(str(tmp_path / "t.json"), device_name="fake", timer=lambda job, algo, tune=0: 1.0)
    assert "direct" not in t.candidates(_job(256, False, 1000))      # 1000 < 256^2/4
    assert "direct" in t.candidates(_job(256, False, 20000))
    assert "direct" not in t.candidates(_job(128, True, 4095))
    assert set(t.candidates(_job(32, True, 256))) == {"marker", "tile", "direct"}


def test_knobs_tuned_one_axis_at_a_time_and_cached(tmp_path):
    """After the algorithm, each knob axis is timed from the built-in setting
    and the best bits are kept; the cache returns them without timing."""
    from cuda_voxelizer_amd import _capi
    from cuda_voxelizer_amd.autotune import Tuner, describe_tune
    seen = []

    def timer(job, algo, tune=0):
        seen.append((algo, tune))
        if algo != "chunk":
            return 5.0
        t = 2.0
        if tune & _capi.tune_waves(3) == _capi.tune_waves(2):
            t -= 0.5   # six waves per SIMD wins
        if tune & (15 << 24) == _capi.tune_big(10):
            t -= 0.25  # the larger queue threshold wins
        return t

    path = str(tmp_path / "k.json")
    job = _job(256, False, 200_000)
    job.tune = 0
    t = Tuner(path, device_name="fake", timer=timer)
    t.tune_jobs([job])
    assert job.algo == "chunk"
    assert job.tune == _capi.tune_waves(2) | _capi.tune_big(10)
    assert describe_tune("chunk", job.tune) == "waves=6,big=1024"
    # the big axis was timed with the waves axis already at its best
    assert ("chunk", _capi.tune_waves(2) | _capi.tune_big(6)) in seen
    n = len(seen)
    job2 = _job(256, False, 200_000)
    job2.tune = 0
    Tuner(path, device_name="fake", timer=timer).tune_jobs([job2])
    assert len(seen) == n and job2.tune == job.tune
    job3 = _job(256, False, 200_000)
    job3.tune = 7
    Tuner(path, device_name="fake", timer=timer).tune_jobs([job3], knobs=False)
    assert job3.algo == "chunk" and job3.tune == 0  # the fixed defaults, for comparison


def test_settle_keeps_the_faster_knobs_for_the_whole_batch(tmp_path):
    """Knobs tuned per class are checked on the whole batch: they stay when the
    batch is faster with them, and go back to the built-in settings otherwise."""
    from cuda_voxelizer_amd.autotune import Tuner
    t = Tuner(str(tmp_path / "t.json"), device_name="fake", timer=lambda job, algo, tune=0: 1.0)
    seen = []

    def make_jobs():
        jobs = [_job(64, False, 1000), _job(128, True, 5000), _job(32, False, 200)]
        for j, tune in zip(jobs, (0x2000000, 0x1000000, 0)):
            j.tune = tune
        return jobs

    def timer_fixed_faster(jobs, tunes):
        seen.append(list(tunes))
        return 2.0 if any(tunes) else 1.5

    jobs = make_jobs()
    assert t.settle(jobs, batch_timer=timer_fixed_faster) is False
    assert [j.tune for j in jobs] == [0, 0, 0]
    tuned = [0x2000000, 0x1000000, 0]
    assert seen == [tuned, [0, 0, 0], [0, 0, 0], tuned]  # alternating, each setting twice

    jobs = make_jobs()
    assert t.settle(jobs, batch_timer=lambda jobs, tunes: 1.0 if any(tunes) else 1.5) is True
    assert [j.tune for j in jobs] == [0x2000000, 0x1000000, 0]

    untuned = [_job(64, False, 1000)]
    untuned[0].tune = 0
    assert t.settle(untuned, batch_timer=lambda jobs, tunes: 1 / 0) is True  # nothing to check, nothing timed
