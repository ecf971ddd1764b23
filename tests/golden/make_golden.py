"""Writes the golden fixtures under tests/golden/ (run in the build container).

Inputs taken from the reference's own data files (read as data, never run):
  * test_models/bunny.OBJ -- the only mesh fixture of the reference; its CI runs
    `cuda_voxelizer -f test_models/bunny.OBJ -s 64 -cpu` (.github/workflows/autobuild.yml:62).
    Stored as float32 vertices + int32 faces (bunny.npz).
  * src/morton_LUTs.h     -- the three 256-entry Morton LUTs.  Only a SHA-256
    digest of the 768 values and a few spot values are stored (morton_luts.json),
    to pin the closed-form LUTs of the oracle and of the kernels.

Expected outputs (bunny64_tables.npz, kat_*.npz) come from the oracle
(oracle/vox_oracle.c) AFTER it passed the pins in tests/test_oracle_pins.py;
the reference's CPU path cannot be compiled in this image (it needs trimesh2
and the CUDA runtime headers), so no table here is reference-generated.

Usage: python tests/golden/make_golden.py [--reference /root/reference]
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def parse_luts(path: str) -> np.ndarray:
    text = open(path).read()
    tables = []
    for name in ("host_morton256_x", "host_morton256_y", "host_morton256_z"):
        m = re.search(name + r"\[256\]\s*=\s*\{([^}]*)\}", text)
        vals = [int(t, 16) for t in re.findall(r"0x[0-9A-Fa-f]+", m.group(1))]
        assert len(vals) == 256, (name, len(vals))
        tables.append(vals)
    return np.array(tables, dtype=np.uint32)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    ap.add_argument("--skip-tables", action="store_true")
    args = ap.parse_args()

    from cuda_voxelizer_amd import meshgen
    v, f = meshgen.read_obj(os.path.join(args.reference, "test_models", "bunny.OBJ"))
    np.savez_compressed(os.path.join(HERE, "bunny.npz"), vertices=v, faces=f)
    print("bunny", v.shape, f.shape)

    luts = parse_luts(os.path.join(args.reference, "src", "morton_LUTs.h"))
    doc = {
        "source": "src/morton_LUTs.h (x table :5-39, y table :41-75, z table :77-111)",
        "sha256_le_u32_xyz": hashlib.sha256(luts.astype("<u4").tobytes()).hexdigest(),
        "spot": {f"{axis}[{i}]": int(luts[k, i]) for k, axis in enumerate("xyz") for i in (1, 2, 3, 128, 255)},
    }
    with open(os.path.join(HERE, "morton_luts.json"), "w") as fh:
        json.dump(doc, fh, indent=1)
    print("luts", doc["sha256_le_u32_xyz"])

    if args.skip_tables:
        return
    from oracle import oracle as O
    info = O.info_for_mesh(v, 64, f.shape[0])
    surf_lin, _ = O.surface(info, (v, f), morton=False, mode=O.MODE_REF)
    surf_mor, _ = O.surface(info, (v, f), morton=True, mode=O.MODE_REF)
    sol_lin, cl, _ = O.solid(info, (v, f), morton=False, mode=O.MODE_REF)
    sol_mor, cl2, _ = O.solid(info, (v, f), morton=True, mode=O.MODE_REF)
    np.savez_compressed(os.path.join(HERE, "bunny64_tables.npz"), info=np.frombuffer(bytes(info), np.uint8),
                        surface_linear=surf_lin, surface_morton=surf_mor, solid_linear=sol_lin,
                        solid_morton=sol_mor, solid_clamped=np.array([cl, cl2]))
    print("bunny64 tables: surface bits", int(np.unpackbits(surf_lin.view(np.uint8)).sum()),
          "solid bits", int(np.unpackbits(sol_lin.view(np.uint8)).sum()), "clamped", cl, cl2)


if __name__ == "__main__":
    main()
