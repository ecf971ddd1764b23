"""Per-class algorithm autotuner for batched launches.

A voxelization job's best algorithm (chunk / marker / tile / direct, all
bit-identical) depends on its grid, layout, mode and size: chunk wins on most
surfaces, coarse ones included (it splits their large triangles over several
workgroups), marker on fine solid meshes and tile on coarse ones, direct on
fine solid meshes at small grids.  ``Tuner.choose`` times
every applicable algorithm once per class (grid, layout, surface/solid, log2
size bucket) on the job it is asked about, keeps the fastest, and remembers it
in a JSON cache keyed by the device name, so later jobs and later processes
reuse the decision.

What is timed is the job's cost inside a batch, not alone: a job of a few
hundred triangles runs a handful of workgroups, so alone it measures launch
latency and favours whichever algorithm spreads one mesh over the most
workgroups, while in a batch the other meshes fill the device and only the
work counts.  Each candidate runs as one graph launch of K copies of the job
(each with its own table; K sized so the copies hold about ``fill_triangles``
triangles, 4..64, and their tables stay under ``max_table_bytes``), giving
the time per copy, and as a launch of one copy, giving the job's critical
path: a batch lasts at least as long as its slowest job, so a candidate whose
single-copy time exceeds ``latency_slack`` times the best one (plus 20 us of
launch noise) is ruled out, and the fastest per copy among the rest wins.

Once the algorithm is chosen, its knobs are tuned the same way, one axis at
a time starting from the built-in setting (``KNOBS``): the chunk kernel's
waves per SIMD (5 or 6, two instantiations) and its queue threshold (voxel
boxes above 64 or 1024 voxels are shared by the workgroup, or the built-in
rule: above 256, or far larger than their wave's mean), the marker kernel's
queue threshold (16, 32, 64 or 256 columns), and the tile algorithm's
tile size (half, built-in, twice).  Every setting gives the same table; the
chosen VOX_TUNE_* bits are cached with the algorithm.  ``Tuner.settle`` then
checks the knobs on the whole batch (tuned against built-in, same algorithms)
and keeps the faster: knobs picked per class can split a mixed batch into more
class kernels than it pays for.

Cache file: $VOX_AUTOTUNE_CACHE, else ~/.cache/cuda_voxelizer_amd/autotune.json.
"""
from __future__ import annotations

import json
import math
import os
import time

SURFACE_ALGOS = ("chunk", "tile", "direct")
SOLID_ALGOS = ("marker", "tile", "direct")


def _knobs():
    from ._capi import tune_big, tune_tile, tune_waves
    # algorithm -> axes, each a list of (label, VOX_TUNE_* bits); 0 is the built-in setting
    return {
        "chunk": [[("waves=5", tune_waves(1)), ("waves=6", tune_waves(2))],
                  [("big=64", tune_big(6)), ("big=built-in", 0), ("big=1024", tune_big(10))]],
        "marker": [[("big=16", tune_big(4)), ("big=32", 0), ("big=64", tune_big(6)), ("big=256", tune_big(8))]],
        "tile": [[("tile=half", tune_tile(1)), ("tile=1", 0), ("tile=twice", tune_tile(3))]],
        "direct": [],
    }


def describe_tune(algo: str, tune: int) -> str:
    """Readable form of a job's knob bits for its algorithm ("-" when it has none)."""
    labels = []
    for axis in _knobs().get(algo, []):
        v = tune & _axis_mask(axis)
        hit = [lab for lab, bits in axis if bits == v]
        labels.append(hit[0] if hit else axis[0][0].split("=")[0] + "=built-in")
    return ",".join(labels) if labels else "-"


def _axis_mask(axis) -> int:
    m = 0
    for _, bits in axis:
        m |= bits
    # the field the axis's bits live in (a byte-aligned nibble or two-bit field)
    for lo, width in ((24, 4), (28, 2), (30, 2)):
        field = ((1 << width) - 1) << lo
        if m & field:
            return field
    return m


def default_cache_path() -> str:
    env = os.environ.get("VOX_AUTOTUNE_CACHE")
    if env:
        return env
    return os.path.join(os.path.expanduser("~"), ".cache", "cuda_voxelizer_amd", "autotune.json")


def class_key(grid: int, morton: bool, solid: bool, n_triangles: int) -> str:
    bucket = int(math.log2(max(1, n_triangles)))
    return f"G{grid}-{'morton' if morton else 'linear'}-{'solid' if solid else 'surface'}-n2^{bucket}"


class Tuner:
    def __init__(self, path: str | None = None, device_name: str | None = None, reps: int = 3, timer=None,
                 fill_triangles: int = 2_000_000, max_table_bytes: int = 512 << 20, latency_slack: float = 2.0,
                 tune_knobs: bool = True):
        self.path = path or default_cache_path()
        self.tune_knobs = tune_knobs
        self.reps = reps
        self.fill_triangles = fill_triangles
        self.max_table_bytes = max_table_bytes
        self.latency_slack = latency_slack
        self._timer = timer  # callable(job, algo, tune) -> seconds or (single, per copy); default: the GPU
        self.device_name = device_name or self._device_name()
        self.table: dict = {}
        self._load()

    @staticmethod
    def _device_name() -> str:
        try:
            import torch
            return torch.cuda.get_device_name(0)
        except Exception:
            return "unknown"

    def _load(self) -> None:
        try:
            with open(self.path) as fh:
                self.table = json.load(fh).get(self.device_name, {})
        except (OSError, ValueError):
            self.table = {}

    def save(self) -> None:
        doc = {}
        try:
            with open(self.path) as fh:
                doc = json.load(fh)
        except (OSError, ValueError):
            doc = {}
        doc[self.device_name] = self.table
        os.makedirs(os.path.dirname(self.path) or ".", exist_ok=True)
        tmp = self.path + ".tmp"
        with open(tmp, "w") as fh:
            json.dump(doc, fh, indent=1, sort_keys=True)
        os.replace(tmp, self.path)

    def copies(self, job) -> int:
        """K: copies of `job` in one timed batch launch."""
        n = max(1, int(job.info.n_triangles))
        k = min(64, max(4, -(-self.fill_triangles // n)))
        table_bytes = max(1, int(job.table.numel()) * job.table.element_size())
        return max(1, min(k, self.max_table_bytes // table_bytes))

    def _time_batch(self, job, algo: str, k: int, tune: int = 0) -> float:
        import torch

        from . import Batch, BatchJob
        tables = [torch.empty_like(job.table) for _ in range(k)]
        copies = [BatchJob(job.info, job.tris, t, morton=job.morton, solid=job.solid, algo=algo, tune=tune)
                  for t in tables]
        batch = Batch(copies, use_graph=True)
        try:
            batch.launch()  # warm (code load, workspace)
            torch.cuda.synchronize()
            best = float("inf")
            for _ in range(self.reps):
                t0 = time.perf_counter()
                batch.launch()
                torch.cuda.synchronize()
                best = min(best, time.perf_counter() - t0)
        finally:
            batch.close()
        return best

    def _time_gpu(self, job, algo: str, tune: int = 0) -> tuple:
        """(single-copy seconds, seconds per copy in a K-copy launch)."""
        k = self.copies(job)
        return self._time_batch(job, algo, 1, tune), self._time_batch(job, algo, k, tune) / k

    @staticmethod
    def candidates(job) -> tuple:
        """Algorithms to time.  direct (one lane per triangle, no work splitting)
        only for fine meshes (at least G^2/4 triangles): on a coarse mesh its
        critical path is the largest triangle's whole box on one lane, which
        copies of the job hide when timed together but which a batch class of
        its own exposes."""
        algos = SOLID_ALGOS if job.solid else SURFACE_ALGOS
        grid = int(job.info.grid)
        if int(job.info.n_triangles) * 4 < grid * grid:
            algos = tuple(a for a in algos if a != "direct")
        return algos

    def pick(self, times: dict) -> str:
        """Fastest per copy among the candidates whose single-copy time is
        within latency_slack of the best.  A plain number is both times."""
        single = {a: (t[0] if isinstance(t, (tuple, list)) else t) for a, t in times.items()}
        per = {a: (t[1] if isinstance(t, (tuple, list)) else t) for a, t in times.items()}
        bound = self.latency_slack * min(single.values()) + 20e-6
        ok = [a for a in times if single[a] <= bound]
        return min(ok, key=per.get)

    def choose(self, job) -> str:
        """Best algorithm for `job` (a BatchJob); times the candidates on a miss."""
        return self.choose_tuned(job)[0]

    def choose_tuned(self, job) -> tuple:
        """(algorithm, VOX_TUNE_* bits) for `job`: the algorithm as `choose`
        picks it, then its knobs one axis at a time from the built-in setting."""
        key = class_key(job.info.grid, job.morton, job.solid, job.info.n_triangles)
        hit = self.table.get(key)
        if hit and ("tune" in hit or not self.tune_knobs):
            return hit["algo"], int(hit.get("tune", 0)) if self.tune_knobs else 0
        timer = self._timer or self._time_gpu
        if hit:
            algo, times = hit["algo"], None
        else:
            times = {a: timer(job, a, 0) for a in self.candidates(job)}
            algo = self.pick(times)
        entry = {"algo": algo, "seconds": hit["seconds"] if hit else
                 {a: list(t) if isinstance(t, tuple) else t for a, t in times.items()}}
        tune = 0
        if self.tune_knobs:
            knob_times = {}
            for axis in _knobs().get(algo, []):
                mask = _axis_mask(axis)
                opts = {}
                for label, bits in axis:
                    t = (tune & ~mask) | bits
                    opts[t] = timer(job, algo, t)
                    knob_times[label] = list(opts[t]) if isinstance(opts[t], tuple) else opts[t]
                tune = self.pick(opts)
            entry["tune"] = tune
            entry["knobs"] = describe_tune(algo, tune)
            entry["knob_seconds"] = knob_times
        self.table[key] = entry
        return algo, tune

    def _time_jobs(self, jobs, tunes) -> float:
        """Seconds of one graph launch of `jobs` with the VOX_TUNE_* bits `tunes`
        (best of `reps`), into scratch tables so the jobs' own are untouched."""
        import torch

        from . import Batch, BatchJob
        copies = [BatchJob(j.info, j.tris, torch.empty_like(j.table), morton=j.morton, solid=j.solid, algo=j.algo,
                           tune=t) for j, t in zip(jobs, tunes)]
        batch = Batch(copies, use_graph=True)
        try:
            batch.launch()  # warm (code load, workspaces)
            torch.cuda.synchronize()
            best = float("inf")
            for _ in range(self.reps):
                t0 = time.perf_counter()
                batch.launch()
                torch.cuda.synchronize()
                best = min(best, time.perf_counter() - t0)
        finally:
            batch.close()
        return best

    def settle(self, jobs, batch_timer=None) -> bool:
        """Whole-batch check of the tuned knobs.  Each class's knobs were timed
        on copies of one job; in the mixed batch they can split the chunk jobs
        into more classes, one kernel each, whose tails leave the GPU idle.  So
        the batch is timed with the tuned knobs and with every knob at its
        built-in setting (same algorithms), twice each in the order tuned,
        fixed, fixed, tuned, and the setting with the better best is kept.  Returns
        True when the tuned knobs stay.  `batch_timer(jobs, tunes)` -> seconds
        replaces the GPU timing."""
        tuned = [int(getattr(j, "tune", 0)) for j in jobs]
        if not any(tuned):
            return True
        timer = batch_timer or self._time_jobs
        fixed = [0] * len(jobs)
        # alternate the two settings and keep each one's best, so that a drift of
        # the device between timings (clocks, caches) favours neither
        t_tuned, t_fixed = timer(jobs, tuned), timer(jobs, fixed)
        t_fixed = min(t_fixed, timer(jobs, fixed))
        t_tuned = min(t_tuned, timer(jobs, tuned))
        keep = t_tuned <= t_fixed
        if not keep:
            for j in jobs:
                j.tune = 0
        return keep

    def tune_jobs(self, jobs, knobs: bool = True) -> list:
        """Set every job's algorithm (and, with `knobs`, its VOX_TUNE_* bits)
        from the tuned table, tuning misses first."""
        for j in jobs:
            j.algo, tune = self.choose_tuned(j)
            if hasattr(j, "tune"):
                j.tune = tune if knobs else 0
        self.save()
        return jobs
