"""Host mesh loading (vox_load_mesh, the trimesh2 read of main.cpp:174-175)
and the command-line option parsing (main.cpp:109-154).  No GPU needed."""
import os
import struct

import numpy as np
import pytest

from tests.conftest import GOLDEN


@pytest.fixture(scope="module")
def bunny():
    g = np.load(os.path.join(GOLDEN, "bunny.npz"))
    return g["vertices"], g["faces"]


def write_obj(path, v, f, ref="plain"):
    with open(path, "w") as fh:
        fh.write("# test mesh\n")
        for p in v:
            fh.write("v %.9g %.9g %.9g\n" % tuple(float(c) for c in p))
        fh.write("vn 0 0 1\n")
        for t in f:
            if ref == "slash":
                fh.write("f %d//1 %d//1 %d//1\n" % tuple(int(i) + 1 for i in t))
            else:
                fh.write("f %d %d %d\n" % tuple(int(i) + 1 for i in t))


@pytest.mark.parametrize("ref", ["plain", "slash"])
def test_obj_round_trip(capi, tmp_path, bunny, ref):
    from cuda_voxelizer_amd import io
    v, f = bunny
    p = str(tmp_path / "bunny.obj")
    write_obj(p, v, f, ref)
    lv, lf = io.load_mesh(p)
    assert lv.dtype == np.float32 and lf.dtype == np.int32
    np.testing.assert_array_equal(lv, v)
    np.testing.assert_array_equal(lf, f)


def test_obj_polygons_and_relative_indices(capi, tmp_path):
    from cuda_voxelizer_amd import io
    p = tmp_path / "quad.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 0.5 0.5 1\n"
                 "f 1 2 3 4\nf -1 -5 -4\n")
    v, f = io.load_mesh(str(p))
    assert v.shape == (5, 3)
    assert f.tolist() == [[0, 1, 2], [0, 2, 3], [4, 0, 1]]


def test_obj_bad_index_is_an_error(capi, tmp_path):
    from cuda_voxelizer_amd import io
    from cuda_voxelizer_amd._capi import VoxError
    p = tmp_path / "bad.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nf 1 2 3\n")
    with pytest.raises(VoxError):
        io.load_mesh(str(p))


def test_off_stl_ply_agree(capi, tmp_path, bunny):
    from cuda_voxelizer_amd import io
    v, f = bunny
    v, f = v[:300], f[np.all(f < 300, axis=1)]
    # OFF
    off = tmp_path / "m.off"
    with open(off, "w") as fh:
        fh.write("OFF\n%d %d 0\n" % (len(v), len(f)))
        for p in v:
            fh.write("%.9g %.9g %.9g\n" % tuple(float(c) for c in p))
        for t in f:
            fh.write("3 %d %d %d\n" % tuple(t))
    ov, of = io.load_mesh(str(off))
    np.testing.assert_array_equal(ov, v)
    np.testing.assert_array_equal(of, f)
    # binary STL: three new vertices per facet
    stl = tmp_path / "m.stl"
    with open(stl, "wb") as fh:
        fh.write(b"\0" * 80 + struct.pack("<I", len(f)))
        for t in f:
            fh.write(struct.pack("<3f", 0, 0, 0) + v[t].astype("<f4").tobytes() + b"\0\0")
    sv, sf = io.load_mesh(str(stl))
    np.testing.assert_array_equal(sv[sf], v[f])
    # ASCII STL
    astl = tmp_path / "a.stl"
    with open(astl, "w") as fh:
        fh.write("solid t\n")
        for t in f:
            fh.write(" facet normal 0 0 1\n  outer loop\n")
            for k in t:
                fh.write("   vertex %.9g %.9g %.9g\n" % tuple(float(c) for c in v[k]))
            fh.write("  endloop\n endfacet\n")
        fh.write("endsolid t\n")
    av, af = io.load_mesh(str(astl))
    np.testing.assert_array_equal(av[af], v[f])
    # PLY, binary little endian with an extra property, and ascii
    ply = tmp_path / "m.ply"
    with open(ply, "wb") as fh:
        fh.write(("ply\nformat binary_little_endian 1.0\nelement vertex %d\nproperty float x\nproperty float y\n"
                  "property float z\nproperty uchar red\nelement face %d\nproperty list uchar int vertex_indices\n"
                  "end_header\n" % (len(v), len(f))).encode())
        for p in v:
            fh.write(p.astype("<f4").tobytes() + b"\x07")
        for t in f:
            fh.write(b"\x03" + t.astype("<i4").tobytes())
    pv, pf = io.load_mesh(str(ply))
    np.testing.assert_array_equal(pv, v)
    np.testing.assert_array_equal(pf, f)
    aply = tmp_path / "a.ply"
    with open(aply, "w") as fh:
        fh.write("ply\nformat ascii 1.0\nelement vertex %d\nproperty float x\nproperty float y\nproperty float z\n"
                 "element face %d\nproperty list uchar int vertex_index\nend_header\n" % (len(v), len(f)))
        for p in v:
            fh.write("%.9g %.9g %.9g\n" % tuple(float(c) for c in p))
        for t in f:
            fh.write("3 %d %d %d\n" % tuple(t))
    qv, qf = io.load_mesh(str(aply))
    np.testing.assert_array_equal(qv, v)
    np.testing.assert_array_equal(qf, f)


def test_unknown_extension(capi, tmp_path):
    from cuda_voxelizer_amd import io
    from cuda_voxelizer_amd._capi import VoxError
    p = tmp_path / "m.3ds"
    p.write_bytes(b"xx")
    with pytest.raises(VoxError):
        io.load_mesh(str(p))


def test_cli_options(tmp_path, capsys):
    from cuda_voxelizer_amd import cli
    p = tmp_path / "m.obj"
    p.write_text("v 0 0 0\n")
    o = cli.parse(["-f", str(p), "-s", "64", "-o", "BinVox", "-solid"])
    assert o == {"file": str(p), "grid": 64, "output": "binvox", "solid": True}
    assert cli.parse(["-f", str(p)])["output"] == "vox" and cli.parse(["-f", str(p)])["grid"] == 256
    for bad in (["-f", str(p), "-o", "stl"], ["-s", "64"], ["-f", str(tmp_path / "missing.obj")],
                ["-f", str(p), "-cpu"]):
        with pytest.raises(SystemExit) as e:
            cli.parse(bad)
        assert e.value.code == 1
    with pytest.raises(SystemExit) as e:
        cli.parse(["-h"])
    assert e.value.code == 0


_PLY_HEAD = ("ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\nproperty float y\nproperty float z\n"
             "{extra}element face 1\nproperty list uchar int vertex_indices\nend_header\n0 0 0\n1 0 0\n0 1 0\n{face}\n")


@pytest.mark.parametrize("name,data", [
    # 3 * nv wraps 64 bits back to a small number
    ("wrap.off", "OFF\n6148914691236517206 1 0\n0 0 0\n1 0 0\n"),
    ("huge_nv.off", "OFF\n9223372036854775807 0 0\n0 0 0\n"),
    # face indices past int32 or not numbers at all
    ("big.ply", _PLY_HEAD.format(extra="", face="3 0 1 3623878656")),
    ("neg.ply", _PLY_HEAD.format(extra="", face="3 0 1 -1")),
    ("nan.ply", _PLY_HEAD.format(extra="", face="3 0 1 nan")),
    ("inf.ply", _PLY_HEAD.format(extra="", face="3 0 1 1e300")),
])
def test_out_of_range_counts_and_indices_are_errors(capi, tmp_path, name, data):
    """Counts and indices beyond every limit give an error: no wrapped bound,
    and no float-to-int conversion of an out-of-range index."""
    from cuda_voxelizer_amd import io
    from cuda_voxelizer_amd._capi import VoxError
    p = tmp_path / name
    p.write_text(data)
    with pytest.raises(VoxError):
        io.load_mesh(str(p))


def test_ply_binary_index_past_int32_is_an_error(capi, tmp_path):
    from cuda_voxelizer_amd import io
    from cuda_voxelizer_amd._capi import VoxError
    p = tmp_path / "b.ply"
    with open(p, "wb") as fh:
        fh.write(b"ply\nformat binary_little_endian 1.0\nelement vertex 3\nproperty float x\nproperty float y\n"
                 b"property float z\nelement face 1\nproperty list uchar uint vertex_indices\nend_header\n")
        fh.write(np.zeros(9, "<f4").tobytes() + b"\x03" + np.array([0, 1, 2**31], "<u4").tobytes())
    with pytest.raises(VoxError):
        io.load_mesh(str(p))


def test_ply_element_without_properties_is_skipped(capi, tmp_path):
    """An element with no properties has no data, whatever its count."""
    from cuda_voxelizer_amd import io
    p = tmp_path / "e.ply"
    p.write_text(_PLY_HEAD.format(extra="element junk 1000000000000000000\n", face="3 0 1 2"))
    v, f = io.load_mesh(str(p))
    assert v.shape == (3, 3) and f.tolist() == [[0, 1, 2]]


def test_ply_binary64_coordinates_round_to_nearest_float(capi, tmp_path):
    """binary64 coordinates round once to binary32, overflow included: past
    FLT_MAX by less than half an ulp (2^103) gives FLT_MAX, from half an ulp
    on infinity (the tie rounds to even), as numpy's IEEE conversion does."""
    from cuda_voxelizer_amd import io
    fmax = float(np.finfo(np.float32).max)
    xyz = np.array([[fmax + 2.0**102, -(fmax + 2.0**102), 1e300],
                    [fmax + 2.0**103, -1e300, 1.0 / 3.0],
                    [0.1, -2.5e-46, 5e-324]], "<f8")
    p = tmp_path / "d.ply"
    with open(p, "wb") as fh:
        fh.write(b"ply\nformat binary_little_endian 1.0\nelement vertex 3\nproperty double x\nproperty double y\n"
                 b"property double z\nelement face 1\nproperty list uchar int vertex_indices\nend_header\n")
        fh.write(xyz.tobytes() + b"\x03" + np.array([0, 1, 2], "<i4").tobytes())
    v, _ = io.load_mesh(str(p))
    with np.errstate(over="ignore"):
        exp = xyz.astype(np.float32)
    assert exp[0, 0] == np.float32(fmax) and np.isinf(exp[1, 0])
    np.testing.assert_array_equal(v, exp)
