"""Per-class algorithm autotuner.

A voxelization job's best algorithm (chunk / tile / direct, all bit-identical)
depends on its grid, layout, mode and size: chunk wins on coherent surface
meshes with little word sharing across chunks, tile on large grids and solid
fills, direct on tiny jobs where one launch of one-triangle-per-lane work is
all there is.  ``Tuner.choose`` times every applicable algorithm once per class
(grid, layout, surface/solid, log2 size bucket) on the job it is asked about,
keeps the fastest, and remembers it in a JSON cache keyed by the device name,
so later jobs and later processes reuse the decision.

Cache file: $VOX_AUTOTUNE_CACHE, else ~/.cache/cuda_voxelizer_amd/autotune.json.
"""
from __future__ import annotations

import json
import math
import os
import time

SURFACE_ALGOS = ("chunk", "tile", "direct")
SOLID_ALGOS = ("tile", "direct")


def default_cache_path() -> str:
    env = os.environ.get("VOX_AUTOTUNE_CACHE")
    if env:
        return env
    return os.path.join(os.path.expanduser("~"), ".cache", "cuda_voxelizer_amd", "autotune.json")


def class_key(grid: int, morton: bool, solid: bool, n_triangles: int) -> str:
    bucket = int(math.log2(max(1, n_triangles)))
    return f"G{grid}-{'morton' if morton else 'linear'}-{'solid' if solid else 'surface'}-n2^{bucket}"


class Tuner:
    def __init__(self, path: str | None = None, device_name: str | None = None, reps: int = 3, timer=None):
        self.path = path or default_cache_path()
        self.reps = reps
        self._timer = timer  # callable(job, algo) -> seconds; default times on the GPU
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

    def _time_gpu(self, job, algo: str) -> float:
        import torch

        from . import voxelize_solid_table, voxelize_surface
        fn = voxelize_solid_table if job.solid else voxelize_surface
        table = torch.empty_like(job.table)
        fn(job.info, job.tris, morton=job.morton, table=table, algo=algo)  # warm (workspace, code load)
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(self.reps):
            t0 = time.perf_counter()
            fn(job.info, job.tris, morton=job.morton, table=table, algo=algo)
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        return best

    def choose(self, job) -> str:
        """Best algorithm for `job` (a BatchJob); times the candidates on a miss."""
        key = class_key(job.info.grid, job.morton, job.solid, job.info.n_triangles)
        hit = self.table.get(key)
        if hit:
            return hit["algo"]
        timer = self._timer or self._time_gpu
        times = {a: timer(job, a) for a in (SOLID_ALGOS if job.solid else SURFACE_ALGOS)}
        algo = min(times, key=times.get)
        self.table[key] = {"algo": algo, "seconds": times}
        return algo

    def tune_jobs(self, jobs) -> list:
        """Set every job's algorithm from the tuned table (tuning misses first)."""
        for j in jobs:
            j.algo = self.choose(j)
        self.save()
        return jobs
