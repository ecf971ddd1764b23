"""Continuous-batching voxelization server.

Requests (a mesh, a grid, surface or solid) arrive at any time.  Each call to
:meth:`VoxServer.step`:

1. collects micro-batches whose completion event has fired and returns their
   results (device-resident tables, with arrival / admission / launch /
   completion times);
2. admits queued requests in arrival order while the HBM budget allows: the
   request's table is allocated and its mesh is uploaded once and kept
   device-resident for later requests of the same mesh (an LRU cache evicts
   idle meshes when the budget is short);
3. launches admitted requests as one micro-batch on a free runner
   (``vox_runner_*``: the batched kernels, descriptors staged through pinned
   memory, no host wait), round-robin over several streams so micro-batches
   overlap on the device while new requests keep arriving.

``serve_plain`` is the comparison path: one plain per-request call
(``vox_surface`` / ``vox_solid``) per request, in arrival order.

Budget accounting counts table bytes and resident mesh bytes (the runners'
binning workspaces are outside it); a request larger than the whole budget is
rejected.
"""
from __future__ import annotations

import collections
import ctypes
import time
from dataclasses import dataclass

import numpy as np

from . import _capi
from ._capi import VOX_LAYOUT_LINEAR, VOX_LAYOUT_MORTON, check

__all__ = ["Request", "Result", "VoxServer", "serve_plain", "latency_stats"]


@dataclass
class Request:
    rid: int
    mesh_key: object            # hashable id of the mesh (device cache key)
    tris: np.ndarray            # host float32 [N, 3, 3]
    grid: int
    solid: bool
    morton: bool = False
    arrival: float = 0.0        # server clock, seconds


@dataclass
class Result:
    rid: int
    table: object               # int32 device tensor
    info: object                # VoxInfo
    arrival: float
    admitted: float
    launched: float
    done: float
    batch_size: int = 1

    @property
    def latency(self) -> float:
        return self.done - self.arrival


@dataclass
class _Admitted:
    req: Request
    info: object
    tris_dev: object
    table: object
    admitted: float
    nbytes: int


@dataclass
class _InFlight:
    jobs: list
    event: object
    launched: float
    runner: int


@dataclass
class _Mesh:
    tensor: object              # float32 device tensor (9 * n_triangles)
    nbytes: int
    users: int                  # admitted or in-flight requests reading it
    ready: object               # event recorded after its upload
    host: object = None         # pinned source, kept until the upload has run


class _Runner:
    def __init__(self, torch, device):
        self.stream = torch.cuda.Stream(device=device)
        h = ctypes.c_void_p()
        check(_capi.load().vox_runner_create(ctypes.c_void_p(self.stream.cuda_stream), ctypes.byref(h)))
        self.h = h
        self.busy = False

    def launch(self, items) -> None:
        arr = (_capi.vox_batch_item * len(items))(*items)
        check(_capi.load().vox_runner_launch(self.h, arr, len(items)))

    def close(self) -> None:
        if self.h:
            _capi.load().vox_runner_destroy(self.h)
            self.h = None


class VoxServer:
    def __init__(self, budget_bytes: int = 64 << 30, streams: int = 2, max_batch: int = 64,
                 device: str = "cuda:0", clock=time.perf_counter):
        import torch
        self.torch = torch
        self.device = torch.device(device)
        self.budget = int(budget_bytes)
        self.max_batch = int(max_batch)
        self.clock = clock
        self.runners = [_Runner(torch, self.device) for _ in range(max(1, streams))]
        self._next_runner = 0
        self.queue: collections.deque[Request] = collections.deque()
        self.admitted: collections.deque[_Admitted] = collections.deque()
        self.inflight: list[_InFlight] = []
        self.copy_stream = torch.cuda.Stream(device=self.device)
        self.meshes: collections.OrderedDict = collections.OrderedDict()  # key -> _Mesh (LRU order)
        self.boxes: dict = {}  # key -> cube bounding box (createMeshBBCube of the mesh bbox)
        self.used = 0
        self.rejected: list[int] = []
        self.stats = {"batches": 0, "jobs": 0, "mesh_uploads": 0, "mesh_hits": 0}

    # ---- public API ------------------------------------------------------------------
    def submit(self, req: Request) -> None:
        self.queue.append(req)

    def step(self) -> list[Result]:
        out = self._collect()
        self._admit()
        self._launch()
        return out

    def preload(self, mesh_key, tris: np.ndarray, box=None) -> None:
        """Make a mesh resident before its requests arrive (counted against the
        budget like any resident mesh; evictable once idle)."""
        if mesh_key in self.meshes:
            return
        req = Request(-1, mesh_key, tris, 0, False)
        self._mesh(req)
        self.meshes[mesh_key].users = 0
        self.used += self._mesh_bytes(req)
        if box is not None:
            self.boxes[mesh_key] = box

    def pending(self) -> int:
        return len(self.queue) + len(self.admitted) + sum(len(f.jobs) for f in self.inflight)

    def release(self, result: Result) -> None:
        """Return a result's table to the budget (the tensor becomes invalid)."""
        if result.table is not None:
            self.used -= result.table.numel() * 4
            result.table = None

    def drain(self, timeout: float = 600.0) -> list[Result]:
        out: list[Result] = []
        t0 = self.clock()
        while self.pending():
            out += self.step()
            if self.clock() - t0 > timeout:
                raise TimeoutError("server did not drain")
        return out

    def close(self) -> None:
        self.torch.cuda.synchronize(self.device)
        for r in self.runners:
            r.close()

    # ---- internals -------------------------------------------------------------------
    def _table_bytes(self, grid: int) -> int:
        return int(_capi.load().vox_table_bytes(grid, grid, grid))

    def _mesh_bytes(self, req: Request) -> int:
        return int(req.tris.size) * 4

    def _mesh(self, req: Request):
        """Device copy of the request's mesh: uploaded once on the copy stream
        and kept resident.  Every launch that reads it waits on its ready
        event, whichever runner it lands on."""
        torch = self.torch
        if req.mesh_key in self.meshes:
            m = self.meshes.pop(req.mesh_key)
            m.users += 1
            self.meshes[req.mesh_key] = m
            self.stats["mesh_hits"] += 1
            return m.tensor
        host = torch.from_numpy(np.ascontiguousarray(req.tris, np.float32).reshape(-1)).pin_memory()
        with torch.cuda.stream(self.copy_stream):
            t = host.to(self.device, non_blocking=True)
            ready = torch.cuda.Event()
            ready.record(self.copy_stream)
        # the pinned source lives until the copy has run (dropped in _collect)
        self.meshes[req.mesh_key] = _Mesh(t, self._mesh_bytes(req), 1, ready, host)
        self.stats["mesh_uploads"] += 1
        return t

    def _evict(self, need: int, keep=None) -> None:
        """Drop idle resident meshes (LRU first) until `need` more bytes fit;
        the mesh `keep` (the one the request being admitted reads) stays."""
        for key in list(self.meshes):
            if self.used + need <= self.budget:
                return
            m = self.meshes[key]
            if m.users == 0 and key != keep:
                del self.meshes[key]
                self.used -= m.nbytes

    def _info(self, req: Request):
        """voxinfo of a request; the mesh's cube box is computed once per mesh."""
        from . import VoxInfo, create_mesh_bbox_cube, mesh_bbox
        box = self.boxes.get(req.mesh_key)
        if box is None:
            box = self.boxes[req.mesh_key] = create_mesh_bbox_cube(mesh_bbox(req.tris.reshape(-1, 3)))
        return VoxInfo(box, (req.grid, req.grid, req.grid), req.tris.shape[0])

    def _admit(self) -> None:
        torch = self.torch
        while self.queue:
            req = self.queue[0]
            need = self._table_bytes(req.grid) + (0 if req.mesh_key in self.meshes else self._mesh_bytes(req))
            if need > self.budget:
                self.queue.popleft()
                self.rejected.append(req.rid)
                continue
            if self.used + need > self.budget:
                self._evict(need, keep=req.mesh_key)
                if self.used + need > self.budget:
                    return  # wait for tables to be released / batches to finish
            self.queue.popleft()
            info = self._info(req)
            fresh = req.mesh_key not in self.meshes
            tris_dev = self._mesh(req)
            table = torch.empty(self._table_bytes(req.grid) // 4, dtype=torch.int32, device=self.device)
            self.used += self._table_bytes(req.grid) + (self._mesh_bytes(req) if fresh else 0)
            self.admitted.append(_Admitted(req, info, tris_dev, table, self.clock(), self._table_bytes(req.grid)))

    def _launch(self) -> None:
        torch = self.torch
        while self.admitted:
            idx = None
            for k in range(len(self.runners)):
                j = (self._next_runner + k) % len(self.runners)
                if not self.runners[j].busy:
                    idx = j
                    break
            if idx is None:
                return
            self._next_runner = (idx + 1) % len(self.runners)
            runner = self.runners[idx]
            jobs = []
            while self.admitted and len(jobs) < self.max_batch:
                jobs.append(self.admitted.popleft())
            items = []
            for a in jobs:
                it = _capi.vox_batch_item()
                ctypes.memmove(ctypes.addressof(it.info), ctypes.addressof(a.info._c), 64)
                it.d_tris = a.tris_dev.data_ptr()
                it.d_table = a.table.data_ptr()
                it.d_offrow = None
                it.layout = VOX_LAYOUT_MORTON if a.req.morton else VOX_LAYOUT_LINEAR
                it.solid = 1 if a.req.solid else 0
                it.flags = 0
                items.append(it)
            # every mesh of the micro-batch must have landed, whichever stream uploaded it
            waited = set()
            for a in jobs:
                m = self.meshes.get(a.req.mesh_key)
                if m is not None and m.ready is not None and a.req.mesh_key not in waited:
                    runner.stream.wait_event(m.ready)
                    waited.add(a.req.mesh_key)
            for a in jobs:  # the caching allocator must not recycle these while the runner uses them
                a.table.record_stream(runner.stream)
                a.tris_dev.record_stream(runner.stream)
            runner.launch(items)
            ev = torch.cuda.Event()
            ev.record(runner.stream)
            runner.busy = True
            self.inflight.append(_InFlight(jobs, ev, self.clock(), idx))
            self.stats["batches"] += 1
            self.stats["jobs"] += len(jobs)

    def _collect(self) -> list[Result]:
        out: list[Result] = []
        still = []
        for f in self.inflight:
            if f.event.query():
                now = self.clock()
                self.runners[f.runner].busy = False
                for a in f.jobs:
                    m = self.meshes.get(a.req.mesh_key)
                    if m is not None:
                        m.users -= 1
                    out.append(Result(a.req.rid, a.table, a.info, a.req.arrival, a.admitted, f.launched, now,
                                      len(f.jobs)))
            else:
                still.append(f)
        self.inflight = still
        for m in self.meshes.values():  # finished uploads no longer need their pinned source
            if m.host is not None and m.ready.query():
                m.host = None
        return out


def serve_plain(requests: list[Request], device: str = "cuda:0", clock=time.perf_counter,
                realtime: bool = True, mesh_cache: dict | None = None,
                box_cache: dict | None = None) -> list[Result]:
    """One plain voxelization call per request, in arrival order (FIFO).

    With ``realtime`` the loop waits for each request's arrival time (relative
    to the start), so latencies include queueing behind earlier requests.
    Meshes (and their cube boxes) are cached per mesh key, as in the server."""
    import torch

    from . import VoxInfo, create_mesh_bbox_cube, mesh_bbox, voxelize_solid_table, voxelize_surface
    dev = torch.device(device)
    cache: dict = {} if mesh_cache is None else mesh_cache
    boxes: dict = {} if box_cache is None else box_cache
    out: list[Result] = []
    t0 = clock()
    for req in sorted(requests, key=lambda r: r.arrival):
        if realtime:
            while clock() - t0 < req.arrival:
                pass
        arrival = t0 + req.arrival if realtime else clock()
        if req.mesh_key not in cache:
            cache[req.mesh_key] = torch.from_numpy(
                np.ascontiguousarray(req.tris, np.float32).reshape(-1)).to(dev)
        box = boxes.get(req.mesh_key)
        if box is None:
            box = boxes[req.mesh_key] = create_mesh_bbox_cube(mesh_bbox(req.tris.reshape(-1, 3)))
        info = VoxInfo(box, (req.grid, req.grid, req.grid), req.tris.shape[0])
        admitted = clock()
        if req.solid:
            table = voxelize_solid_table(info, cache[req.mesh_key], morton=req.morton)
        else:
            table = voxelize_surface(info, cache[req.mesh_key], morton=req.morton)
        torch.cuda.synchronize(dev)
        done = clock()
        out.append(Result(req.rid, table, info, arrival, admitted, admitted, done))
    return out


def latency_stats(results: list[Result], span: float | None = None) -> dict:
    lat = np.array([r.latency for r in results]) * 1e3
    if not len(lat):
        return {"n": 0}
    if span is None:
        span = max(r.done for r in results) - min(r.arrival for r in results)
    return {"n": int(len(lat)), "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "grids_per_s": len(lat) / span if span > 0 else float("nan")}
