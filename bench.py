"""Voxelization throughput on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--config torus512m|...] [--algo tile|direct]

A step is one voxelization of the whole workload (the table is fully defined
by the call: zero-fill fused, mesh already resident in HBM).  Default
workload: the 512^3 Morton surface voxelization of a jittered 1024x512-quad
torus (1,048,576 FP32 triangles, 10% slivers).  Prints ONE JSON line.

Roofline: algorithmic bytes per step B = 36*N_tri + G^3/8 (read every triangle
once, write the table once) over the step's device time, measured with HIP
events on the stream the library launches on, against 8 TB/s (MI355X HBM3E).
The CPU baseline is the oracle's reference-structured OpenMP path (one global
critical section per bit, cpu_voxelizer.cpp:7-15,53) on this host's cores; the
same leg checks the GPU table bit-for-bit (mismatches must be 0).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

CONFIGS = {
    # name: (builder, grid, morton, solid, description)
    "torus512m": ("torus", 512, True, False, "torus 1024x512 quads + 10% slivers, 512^3 Morton surface"),
    "torus1024": ("torus1024", 1024, False, False, "torus 1024x512 quads + 10% slivers, 1024^3 linear surface"),
    "ico64": ("ico4", 64, False, False, "icosphere L4 +-2% jitter, 64^3 linear surface"),
    "sh1024s": ("sh8", 1024, False, True, "SH-displaced icosphere L8, 1024^3 linear solid"),
}


def build_mesh(kind: str):
    from cuda_voxelizer_amd import meshgen
    if kind == "torus":
        return meshgen.torus_slivers(grid=512)
    if kind == "torus1024":
        return meshgen.torus_slivers(grid=1024)
    if kind == "ico4":
        v, f = meshgen.icosphere_jitter()
        return meshgen.to_aos(v, f)
    if kind == "sh8":
        v, f = meshgen.icosphere_sh()
        return meshgen.to_aos(v, f)
    raise ValueError(kind)


def load_traffic(workload: str):
    p = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if os.path.exists(p):
        with open(p) as fh:
            return json.load(fh).get("hbm_bytes_per_launch")
    return None


def cpu_baseline_leg(tris, info, grid, morton, solid, gpu_table, budget_s=20.0):
    """Times the oracle's reference-structured path on a bounded sample and
    checks the GPU table against the oracle (fast mode) on the full workload."""
    from oracle import oracle as O
    cores = os.cpu_count() or 1
    cores = min(cores, 16)
    O.set_threads(cores)
    oinfo = O.info_from_bytes(info.as_bytes())
    # bit check on the full workload
    if solid:
        exp, _, _ = O.solid(oinfo, tris, morton=morton, mode=O.MODE_FAST)
    else:
        exp, _ = O.surface(oinfo, tris, morton=morton, mode=O.MODE_FAST)
    mismatches = int(np.unpackbits((exp ^ gpu_table).view(np.uint8)).sum())
    # timed reference-structured run: grow the sample until ~budget_s of work
    n = tris.shape[0]
    m = min(n, 65536)
    while True:
        if solid:
            _, _, ms = O.solid(oinfo, tris[:m], morton=morton, mode=O.MODE_REF)
        else:
            _, ms = O.surface(oinfo, tris[:m], morton=morton, mode=O.MODE_REF)
        if m == n or ms > 1000.0 * budget_s / 4:
            break
        m = min(n, m * 4)
    return {
        "value": m / (ms / 1e3),
        "unit": "triangles/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{m} of {n} triangles ({'solid' if solid else 'surface'}, {grid}^3, "
                  f"{'morton' if morton else 'linear'}), omp critical per bit, {ms:.1f} ms",
    }, mismatches


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="torus512m", choices=sorted(CONFIGS))
    ap.add_argument("--algo", default="auto", choices=["auto", "tile", "chunk", "direct"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the voxelizer shards by z-slab with no collective; see the sharding design)", file=sys.stderr)
        sys.exit(2)

    import torch
    import cuda_voxelizer_amd as V

    kind, grid, morton, solid, desc = CONFIGS[args.config]
    tris = build_mesh(kind)
    n = tris.shape[0]
    dev = torch.device("cuda:0")
    info = V.voxinfo_for_mesh(tris.reshape(-1, 3), grid, n)
    d_tris = torch.from_numpy(tris.reshape(-1)).to(dev)
    table = torch.empty(V.table_words(grid), dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()

    def step():
        if solid:
            V.voxelize_solid_table(info, d_tris, morton=morton, table=table, algo=args.algo, stream=stream)
        else:
            V.voxelize_surface(info, d_tris, morton=morton, table=table, algo=args.algo, stream=stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        step()
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    launch_ms = float(np.mean(ev_ms))
    gpu_table = table.cpu().numpy().view(np.uint32).copy()

    alg_bytes = 36 * n + (grid ** 3) // 8
    achieved = alg_bytes / (launch_ms * 1e-3) / 1e9
    tri_s = n * args.steps / wall
    workload = args.config
    out = {
        "metric": "triangles/s at zero bit mismatches",
        "value": tri_s,
        "unit": "triangles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": desc, "name": workload, "n_triangles": n, "grid": grid,
                   "layout": "morton" if morton else "linear", "mode": "solid" if solid else "surface",
                   "algo": args.algo},
        "grids_per_s": args.steps / wall,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": load_traffic(workload),
                     "algorithmic_bytes": alg_bytes, "launch_ms": launch_ms,
                     "launch_ms_p50": float(np.median(ev_ms))},
    }
    if not args.no_cpu_baseline:
        cpu, mism = cpu_baseline_leg(tris, info, grid, morton, solid, gpu_table)
        out["cpu_baseline"] = cpu
        out["mismatches"] = mism
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
