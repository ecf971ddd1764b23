"""Command line of the voxelizer, with the reference program's options
(main.cpp:109-154) and steps (main.cpp:158-262):

  python -m cuda_voxelizer_amd -f <mesh> [-s <grid>] [-o vox|binvox|obj|obj_points|morton] [-solid]

read mesh (OBJ / OFF / STL / PLY) -> bounding box of all vertices -> cube
padding (createMeshBBCube) -> voxinfo -> triangles packed and voxelized on
the GPU -> the output file written by the device encoders.  Defaults as in
the reference: grid 256, output ``vox``.  Output files are named from the
input path including its extension, as the reference does (``bunny.obj`` ->
``bunny.obj_256.binvox``; Morton: ``bunny.obj.bin``).

Differences: ``-cpu`` (the reference's CPU voxelizer) is refused -- this build
voxelizes on the GPU only, and a missing GPU is an error, not a fallback.
"""
from __future__ import annotations

import sys
import time

VERSION = "cuda_voxelizer_amd 0.1 (MI355X)"
OUTPUTS = {"binvox": "binvox file", "morton": "morton encoded blob", "obj_points": "obj file (pointcloud)",
           "obj": "obj file (cubes)", "vox": "magicavoxel file"}


def print_help(out=sys.stdout) -> None:
    print("\n## HELP  \nProgram options: \n", file=out)
    print(" -f <path to model file: .ply, .obj, .off, .stl> (required)", file=out)
    print(" -s <voxelization grid size, power of 2: 8 -> 512, 1024, ... (default: 256)>", file=out)
    print(" -o <output format: vox, binvox, obj, obj_points or morton (default: vox)>", file=out)
    print(" -solid : Force solid voxelization (experimental, needs watertight model)\n", file=out)
    print("Example: python -m cuda_voxelizer_amd -f /home/user/bunny.ply -s 512\n", file=out)


def parse(argv: list[str]) -> dict:
    """The reference's parseProgramParameters: returns options or raises SystemExit."""
    if len(argv) < 1:
        print("Not enough program parameters. \n ")
        print_help()
        raise SystemExit(0)
    opts = {"file": None, "grid": 256, "output": "vox", "solid": False}
    i = 0
    while i < len(argv):
        a = argv[i]
        if a == "-f" and i + 1 < len(argv):
            opts["file"] = argv[i + 1]
            i += 1
        elif a == "-s" and i + 1 < len(argv):
            try:
                opts["grid"] = int(argv[i + 1])
            except ValueError:
                print(f"[Err] Grid size is not an integer: {argv[i + 1]}")
                raise SystemExit(1)
            i += 1
        elif a == "-h":
            print_help()
            raise SystemExit(0)
        elif a == "-o" and i + 1 < len(argv):
            o = argv[i + 1].lower()
            if o not in OUTPUTS:
                print(f"[Err] Unrecognized output format: {o}, valid options are binvox (default), morton, obj, "
                      "obj_points or vox")
                raise SystemExit(1)
            opts["output"] = o
            i += 1
        elif a == "-cpu":
            print("[Err] -cpu: CPU voxelization is not part of this build; it voxelizes on the GPU only.")
            raise SystemExit(1)
        elif a == "-solid":
            opts["solid"] = True
        i += 1
    if not opts["file"]:
        print("[Err] You didn't specify a file using -f (path). This is required. Exiting. ")
        print("Example: python -m cuda_voxelizer_amd -f /home/user/bunny.ply -s 512")
        raise SystemExit(1)
    import os
    if not os.path.isfile(opts["file"]):
        print(f"[Err] File does not exist / cannot access: {opts['file']} ")
        raise SystemExit(1)
    if opts["grid"] < 1:
        print("[Err] Grid size must be positive")
        raise SystemExit(1)
    return opts


def main(argv: list[str] | None = None) -> int:
    t0 = time.perf_counter()
    argv = sys.argv[1:] if argv is None else argv
    print("## CUDA VOXELIZER (MI355X build)")
    print(VERSION)
    print("\n## PROGRAM PARAMETERS ")
    opts = parse(argv)
    G, fmt = opts["grid"], opts["output"]
    print(f"[Info] Filename: {opts['file']} ")
    print(f"[Info] Grid size: {G} ")
    print(f"[Info] Output format: {OUTPUTS[fmt]} ")
    print(f"[Info] Using Solid Voxelization: {'Yes' if opts['solid'] else 'No'} (default: No)")

    import torch

    import cuda_voxelizer_amd as V
    from cuda_voxelizer_amd import io
    if not torch.cuda.is_available():
        print("[Err] No HIP device visible: this build needs a GPU.")
        return 1
    print("\n## READ MESH ")
    print(f"[I/O] Reading mesh from {opts['file']} ")
    v, f = io.load_mesh(opts["file"])
    print(f"[Mesh] Number of triangles: {f.shape[0]} ")
    print(f"[Mesh] Number of vertices: {v.shape[0]} ")
    if v.shape[0] == 0:
        print("[Err] The mesh has no vertices.")
        return 1
    morton = fmt == "morton"
    if morton and (G & (G - 1)):
        print("[Err] Morton output needs a power-of-two grid size.")
        return 1
    print("\n## VOXELISATION SETUP ")
    info, tris = V.prepare_mesh(v, f, G)
    print(f"[Voxelization] Bounding Box: {list(info.bbox.min)} to {list(info.bbox.max)} ")
    print(f"[Voxelization] Grid size: {G} {G} {G} ")
    print(f"[Voxelization] Triangles: {info.n_triangles} ")
    print(f"[Voxelization] Unit length: x: {info.unit[0]:g} y: {info.unit[1]:g} z: {info.unit[2]:g}")
    print("\n## GPU VOXELISATION ")
    t1 = time.perf_counter()
    if opts["solid"]:
        table = V.voxelize_solid_table(info, tris, morton=morton)
    else:
        table = V.voxelize_surface(info, tris, morton=morton)
    torch.cuda.synchronize()
    print(f"[Perf] Voxelization GPU time: {(time.perf_counter() - t1) * 1e3:.1f} ms")
    print("\n## FILE OUTPUT ")
    name = opts["file"]
    if morton:
        path = io.write_binary(table, name)
    elif fmt == "binvox":
        path = io.write_binvox(table, info, name)
    elif fmt == "obj_points":
        path = io.write_obj_pointcloud(table, info, name)
    elif fmt == "obj":
        path = io.write_obj_cubes(table, info, name)
    else:
        path = io.write_vox(table, info, name)
    print(f"[I/O] Wrote {path}")
    print("\n## STATS ")
    print(f"[Perf] Total runtime: {(time.perf_counter() - t0) * 1e3:.1f} ms ")
    return 0


if __name__ == "__main__":
    sys.exit(main())
