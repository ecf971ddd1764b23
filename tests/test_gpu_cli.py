"""The command line end to end on the reference's CI case (bunny, -s 64,
autobuild.yml:62): every output file equals the oracle writers applied to the
oracle's table of the same mesh."""
import os

import numpy as np
import pytest

from tests.conftest import GOLDEN
from tests.test_mesh_io import write_obj

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bunny_obj(tmp_path_factory):
    g = np.load(os.path.join(GOLDEN, "bunny.npz"))
    p = str(tmp_path_factory.mktemp("cli") / "bunny.OBJ")
    write_obj(p, g["vertices"], g["faces"])
    return p


@pytest.fixture(scope="module")
def golden():
    return np.load(os.path.join(GOLDEN, "bunny64_tables.npz"))


def binvox_header(info_bytes: bytes) -> bytes:
    a = np.frombuffer(info_bytes, np.uint8)
    mn, mx = a[0:12].view(np.float32), a[12:24].view(np.float32)
    ext = [np.float32(mx[k] - mn[k]) for k in range(3)]
    s = ext[0] if not ext[0] < ext[1] else ext[1]
    s = s if not s < ext[2] else ext[2]
    return ("#binvox 1\ndim 64 64 64\ntranslate %g %g %g\nscale %g\ndata\n" %
            (float(mn[0]), float(mn[1]), float(mn[2]), float(s))).encode()


@pytest.mark.parametrize("fmt,solid", [("binvox", False), ("vox", False), ("obj_points", False), ("obj", False),
                                       ("morton", False), ("binvox", True), ("vox", True)])
def test_cli_outputs_match_oracle(oracle, gpu, bunny_obj, golden, fmt, solid):
    from cuda_voxelizer_amd import cli
    args = ["-f", bunny_obj, "-s", "64", "-o", fmt] + (["-solid"] if solid else [])
    print("cli", args, flush=True)
    assert cli.main(args) == 0
    kind = "solid" if solid else "surface"
    lin = golden[f"{kind}_linear"]
    if fmt == "morton":
        path, exp = bunny_obj + ".bin", golden[f"{kind}_morton"].tobytes()
    elif fmt == "binvox":
        path = bunny_obj + "_64.binvox"
        exp = binvox_header(golden["info"].tobytes()) + oracle.binvox_body(lin, 64)
    elif fmt == "vox":
        path, exp = bunny_obj + "_64.vox", oracle.vox_file(lin, 64)
    elif fmt == "obj_points":
        path, exp = bunny_obj + "_64_pointcloud.obj", oracle.obj_points(lin, 64)
    else:
        path, exp = bunny_obj + "_64_voxels.obj", oracle.obj_cubes(lin, 64)
    with open(path, "rb") as fh:
        assert fh.read() == exp
