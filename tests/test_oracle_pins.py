"""Pins of the CPU oracle (oracle/vox_oracle.c) against what fixes each of its
functions independently of its own code.

The reference's CPU path cannot be compiled in this image (it needs trimesh2
and the CUDA runtime headers, neither present), so the pins are:
  * the reference's data files: the Morton LUTs (digest of morton_LUTs.h) and
    the bunny mesh (its only test fixture, autobuild.yml:62);
  * known answers stated by the reference (util.h:75-78 createMeshBBCube example);
  * closed forms (bit interleave, exact power-of-two division, MSB-first layout);
  * independent float64 brute force (tests/bruteforce.py: 13-axis SAT for the
    surface test, +x ray parity with truncation for the solid fill), away from ties;
  * hand-derived known answers for degenerate and axis-aligned triangles
    (the solid thresholds, where a float compare in place of the reference's
    double one changes the answer, are pinned in test_oracle_thresholds.py).
Each check is chosen so that a plausible slip in the oracle (swapped axis,
dropped term, wrong sign, wrong bit order) fails it.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from tests.bruteforce import solid_parity, surface_sat
from tests.conftest import GOLDEN


def _translated(O, info, tris):
    """vertex - bbox.min in binary32, as the reference's prepass computes it."""
    bmin = np.array(info.bmin[:], np.float32)
    return (tris.astype(np.float32) - bmin).astype(np.float64)


def test_morton_luts_match_reference_digest(oracle):
    doc = json.load(open(os.path.join(GOLDEN, "morton_luts.json")))
    luts = oracle.morton_luts()
    assert hashlib.sha256(luts.astype("<u4").tobytes()).hexdigest() == doc["sha256_le_u32_xyz"]
    for key, val in doc["spot"].items():
        axis, idx = key[0], int(key[2:-1])
        assert int(luts["xyz".index(axis), idx]) == val, key


def _interleave(x, y, z, bits=16):
    c = 0
    for b in range(bits):
        c |= ((x >> b) & 1) << (3 * b) | ((y >> b) & 1) << (3 * b + 1) | ((z >> b) & 1) << (3 * b + 2)
    return c


def test_morton_encode_is_bit_interleave(oracle):
    rng = np.random.default_rng(1)
    for _ in range(2000):
        x, y, z = (int(v) for v in rng.integers(0, 1 << 16, 3))
        assert oracle.morton_encode(x, y, z) == _interleave(x, y, z)
    # the reference's byte-2 term is shifted out: coordinates are taken mod 2^16
    assert oracle.morton_encode(70000, 5, 6) == _interleave(70000 - 65536, 5, 6)
    # axis order: x -> bit 0, y -> bit 1, z -> bit 2
    assert [oracle.morton_encode(*p) for p in ((1, 0, 0), (0, 1, 0), (0, 0, 1))] == [1, 2, 4]


def test_bbox_cube_reference_example(oracle):
    # util.h:75-78: (1,2,3)-(4,4,4) becomes the cube (1,1.5,2)-(4,4.5,5) before the 1/10001 padding
    mn, mx = oracle.bbox_cube([1, 2, 3], [4, 4, 4])
    eps = np.float32(3.0) / np.float32(10001.0)
    np.testing.assert_array_equal(mn, np.array([1, 1.5, 2], np.float32) - eps)
    np.testing.assert_array_equal(mx, np.array([4, 4.5, 5], np.float32) + eps)
    # cube property and centre preservation on random boxes
    rng = np.random.default_rng(2)
    for _ in range(200):
        a = rng.normal(size=3).astype(np.float32)
        b = a + rng.uniform(0.01, 3, size=3).astype(np.float32)
        mn, mx = oracle.bbox_cube(a, b)
        side = (mx - mn).astype(np.float64)
        assert np.ptp(side) <= 4e-6 * side.max()
        np.testing.assert_allclose((mn + mx) / 2, (a + b) / 2, rtol=0, atol=2e-6 * side.max())
        assert side.max() == pytest.approx(float((b - a).max()) * (1 + 2 / 10001), rel=1e-6)


def test_make_info_unit_power_of_two_exact(oracle):
    rng = np.random.default_rng(3)
    for g in (1, 2, 64, 512, 1024):
        a = rng.normal(size=3).astype(np.float32)
        b = a + rng.uniform(0.5, 2, size=3).astype(np.float32)
        info = oracle.make_info(a, b, g, 7)
        # division by a power of two is exact: unit * G == max - min bit-for-bit
        np.testing.assert_array_equal(np.array(info.unit[:], np.float32) * np.float32(g), (b - a).astype(np.float32))
        assert list(info.grid) == [g, g, g] and info.n_triangles == 7


def test_mesh_bbox_is_componentwise_minmax(oracle):
    v = np.random.default_rng(4).normal(size=(1000, 3)).astype(np.float32)
    mn, mx = oracle.mesh_bbox(v)
    np.testing.assert_array_equal(mn, v.min(0))
    np.testing.assert_array_equal(mx, v.max(0))


def _random_tris(rng, n, scale_lo, scale_hi):
    c = rng.uniform(0.1, 0.9, size=(n, 1, 3))
    d = rng.normal(size=(n, 3, 3))
    d /= np.linalg.norm(d, axis=2, keepdims=True)
    s = np.exp(rng.uniform(np.log(scale_lo), np.log(scale_hi), size=(n, 3, 1)))
    return (c + d * s).astype(np.float32)


@pytest.mark.parametrize("grid,scale", [(16, (0.01, 0.05)), (24, (0.02, 0.3)), (32, (0.002, 0.02))])
def test_surface_matches_sat_bruteforce(oracle, grid, scale):
    rng = np.random.default_rng(grid)
    tris = _random_tris(rng, 150, *scale)
    info = oracle.info_for_mesh(tris.reshape(-1, 3), grid, tris.shape[0])
    table, _ = oracle.surface(info, tris)
    occ = oracle.occupancy(table, grid)
    ref, tie = surface_sat(_translated(oracle, info, tris), np.array(info.unit[:], np.float64), grid)
    decided = ~tie
    assert decided.sum() > 0.99 * decided.size
    mism = (occ != ref) & decided
    assert int(mism.sum()) == 0, f"{int(mism.sum())} non-tie voxel(s) differ from SAT"
    assert ref.sum() > 200  # the check is not vacuous


def test_surface_point_triangle_sets_its_voxel(oracle):
    # all three vertices equal: zero-area -> NaN normal -> the plane test never
    # rejects, every edge normal is (0,0) with offset 0 -> all tests pass, so
    # exactly the voxel containing the point is set.
    grid = 8
    info = oracle.make_info([0, 0, 0], [8, 8, 8], grid, 1)
    rng = np.random.default_rng(5)
    for _ in range(50):
        p = rng.uniform(0.05, 7.95, size=3).astype(np.float32)
        tris = np.repeat(p[None, None, :], 3, axis=1)
        table, _ = oracle.surface(info, tris)
        occ = oracle.occupancy(table, grid)
        assert occ.sum() == 1
        z, y, x = np.argwhere(occ)[0]
        assert (x, y, z) == tuple(int(c) for c in np.floor(p))


def test_surface_segment_triangle_matches_box_segment_overlap(oracle):
    # exactly collinear vertices (v1 - v0 == v2 - v1 in binary32): cross(e0,e1)
    # is exactly 0, the normal is NaN and the plane test never rejects; the
    # projections' edge tests then reduce to "the box touches the segment's
    # line", so inside the voxel range the result is the 3D box/segment overlap.
    grid = 16
    info = oracle.make_info([0, 0, 0], [16, 16, 16], grid, 1)
    rng = np.random.default_rng(6)
    checked = 0
    for _ in range(40):
        a = rng.integers(64, 15 * 64, 3) / 64.0
        d = rng.integers(-3 * 64, 3 * 64, 3) / 128.0
        d = np.where(np.abs(a + 2 * d - 8) > 7.5, -d, d)
        tri = np.stack([a, a + d, a + 2 * d]).astype(np.float32)
        assert ((tri[1] - tri[0]) == (tri[2] - tri[1])).all()
        table, _ = oracle.surface(info, tri[None])
        occ = oracle.occupancy(table, grid)
        # brute force: slab clipping of the segment against each voxel box
        p, q = tri[0].astype(np.float64), tri[2].astype(np.float64)
        lo = np.trunc(np.minimum(p, q)).astype(int)
        hi = np.trunc(np.maximum(p, q)).astype(int)
        for z in range(lo[2], hi[2] + 1):
            for y in range(lo[1], hi[1] + 1):
                for x in range(lo[0], hi[0] + 1):
                    bmin = np.array([x, y, z], float)
                    t0, t1, tie = 0.0, 1.0, False
                    for k in range(3):
                        d = q[k] - p[k]
                        if abs(d) < 1e-12:
                            if not (bmin[k] <= p[k] <= bmin[k] + 1):
                                t0, t1 = 1, 0
                            continue
                        ta, tb = (bmin[k] - p[k]) / d, (bmin[k] + 1 - p[k]) / d
                        t0, t1 = max(t0, min(ta, tb)), min(t1, max(ta, tb))
                    gap = t1 - t0
                    if abs(gap) * np.linalg.norm(q - p) < 1e-4:
                        continue
                    assert occ[z, y, x] == (gap > 0), (x, y, z)
                    checked += 1
    assert checked > 100


def test_surface_axis_aligned_triangle_on_voxel_face(oracle):
    # a triangle in the plane z = 3 (exactly a voxel face): the voxel range is
    # trunc(3/1) = 3 only, the plane test gives (0)*(x) = 0 -> not rejected, so
    # layer z=3 is set and layer z=2 (which also touches the plane) is not.
    grid = 8
    info = oracle.make_info([0, 0, 0], [8, 8, 8], grid, 1)
    tri = np.array([[[1.2, 1.3, 3.0], [5.7, 1.4, 3.0], [1.5, 5.6, 3.0]]], np.float32)
    table, _ = oracle.surface(info, tri)
    occ = oracle.occupancy(table, grid)
    assert occ[3].sum() > 0 and occ[2].sum() == 0 and occ.sum() == occ[3].sum()
    # in-plane coverage equals the 2-D triangle/square overlap (closed squares)
    for y in range(8):
        for x in range(8):
            corners = [(x, y), (x + 1, y), (x, y + 1), (x + 1, y + 1)]
            tri2 = tri[0, :, :2].astype(np.float64)
            # SAT in 2-D
            sep = False
            pts = tri2
            for axis in (np.array([1.0, 0]), np.array([0, 1.0])):
                tp, cp = pts @ axis, np.array(corners, float) @ axis
                if tp.max() < cp.min() or cp.max() < tp.min():
                    sep = True
            for i in range(3):
                e = pts[(i + 1) % 3] - pts[i]
                nrm = np.array([-e[1], e[0]])
                tp, cp = pts @ nrm, np.array(corners, float) @ nrm
                if tp.max() < cp.min() - 1e-9 or cp.max() < tp.min() - 1e-9:
                    sep = True
            in_range = (x <= 5 and y <= 5 and x >= 1 and y >= 1)
            assert occ[3, y, x] == (in_range and not sep), (x, y)


@pytest.mark.parametrize("level,grid", [(2, 16), (3, 24), (3, 32)])
def test_solid_matches_ray_parity(oracle, level, grid):
    from cuda_voxelizer_amd import meshgen
    v, f = meshgen.icosphere_jitter(level=level, amp=0.1, seed=level * 100 + grid)
    tris = meshgen.to_aos(v, f)
    info = oracle.info_for_mesh(v, grid, f.shape[0])
    table, clamped, _ = oracle.solid(info, (v, f))
    assert clamped == 0
    occ = oracle.occupancy(table, grid)
    ref, tiecol = solid_parity(_translated(oracle, info, tris), np.array(info.unit[:], np.float64), grid)
    ok = ~tiecol
    assert ok.sum() > 0.97 * ok.size
    mism = (occ != ref) & ok[:, :, None]
    assert int(mism.sum()) == 0
    assert ref.sum() > 100


def test_solid_single_triangle_closed_form(oracle):
    # triangle in the plane x = 5.3 (grid 8, unit 1): every column whose centre is
    # strictly inside the YZ triangle gets bits x = 0 .. int(5.3 - 0.5) = 4.
    grid = 8
    info = oracle.make_info([0, 0, 0], [8, 8, 8], grid, 1)
    tri = np.array([[[5.3, 0.2, 0.3], [5.3, 7.1, 0.6], [5.3, 0.9, 7.4]]], np.float32)
    table, clamped, _ = oracle.solid(info, tri)
    occ = oracle.occupancy(table, grid)
    assert clamped == 0
    A, B, C = tri[0, :, 1:].astype(np.float64)
    for z in range(grid):
        for y in range(grid):
            P = np.array([y + 0.5, z + 0.5])
            def cr(u, v):
                return u[0] * v[1] - u[1] * v[0]

            w = [cr(B - A, P - A), cr(C - B, P - B), cr(A - C, P - C)]
            inside = all(x > 1e-9 for x in w) or all(x < -1e-9 for x in w)
            expect = np.zeros(grid, bool)
            if inside:
                expect[:5] = True
            assert (occ[z, y] == expect).all(), (y, z)


def test_solid_parallel_to_x_is_skipped(oracle):
    # n.x = 0 (a triangle containing the x direction) contributes nothing; the
    # values either side of 1e-6 are in test_oracle_thresholds.py
    grid = 8
    info = oracle.make_info([0, 0, 0], [8, 8, 8], grid, 1)
    tri = np.array([[[0.5, 1.2, 1.3], [7.5, 1.2, 1.3], [3.0, 6.2, 5.9]]], np.float32)
    table, _, _ = oracle.solid(info, tri)
    assert not table.any()


def test_ref_and_fast_modes_agree(oracle):
    from cuda_voxelizer_amd import meshgen
    v, f = meshgen.icosphere_jitter(level=3, seed=9)
    info = oracle.info_for_mesh(v, 32, f.shape[0])
    for morton in (False, True):
        a, _ = oracle.surface(info, (v, f), morton=morton, mode=oracle.MODE_REF)
        b, _ = oracle.surface(info, (v, f), morton=morton, mode=oracle.MODE_FAST)
        assert (a == b).all()
        c, _, _ = oracle.solid(info, (v, f), morton=morton, mode=oracle.MODE_REF)
        d, _, _ = oracle.solid(info, (v, f), morton=morton, mode=oracle.MODE_FAST)
        assert (c == d).all()


def test_layouts_hold_the_same_voxels(oracle):
    # linear and Morton tables decode to the same occupancy; bit order is MSB-first
    from cuda_voxelizer_amd import meshgen
    v, f = meshgen.icosphere_jitter(level=3, seed=11)
    info = oracle.info_for_mesh(v, 32, f.shape[0])
    lin, _ = oracle.surface(info, (v, f))
    mor, _ = oracle.surface(info, (v, f), morton=True)
    np.testing.assert_array_equal(oracle.occupancy(lin, 32), oracle.occupancy(mor, 32, morton=True))
    one = np.array([[[0.5, 0.5, 0.5]] * 3], np.float32)  # point triangle in voxel (0,0,0)
    info1 = oracle.make_info([0, 0, 0], [32, 32, 32], 32, 1)
    t, _ = oracle.surface(info1, one)
    assert t[0] == 0x80000000 and not t[1:].any()


def test_accumulate_semantics(oracle):
    # the reference ORs (surface) / XORs (solid) into the caller's table
    from cuda_voxelizer_amd import meshgen
    v, f = meshgen.icosphere_jitter(level=2, seed=12)
    info = oracle.info_for_mesh(v, 16, f.shape[0])
    base = np.random.default_rng(0).integers(0, 2**32, oracle.table_words(16), dtype=np.uint64).astype(np.uint32)
    s0, _ = oracle.surface(info, (v, f))
    s1, _ = oracle.surface(info, (v, f), table=base.copy())
    assert (s1 == (base | s0)).all()
    d0, _, _ = oracle.solid(info, (v, f))
    d1, _, _ = oracle.solid(info, (v, f), table=base.copy())
    assert (d1 == (base ^ d0)).all()


def test_bunny_fixture_invariants(oracle):
    # the reference's CI case (bunny.OBJ, -s 64 -cpu): every vertex's voxel is
    # set, and the solid fill is contained in the surface's closure.
    g = np.load(os.path.join(GOLDEN, "bunny.npz"))
    v, f = g["vertices"], g["faces"]
    assert v.shape == (2557, 3) and f.shape == (5110, 3)
    info = oracle.info_for_mesh(v, 64, f.shape[0])
    table, _ = oracle.surface(info, (v, f), mode=oracle.MODE_REF)
    occ = oracle.occupancy(table, 64)
    tv = (v - np.array(info.bmin[:], np.float32)) / np.array(info.unit[:], np.float32)
    vox = np.clip(np.trunc(tv).astype(int), 0, 63)
    frac = tv - np.trunc(tv)
    interior = ((frac > 1e-3) & (frac < 1 - 1e-3)).all(1)
    assert occ[vox[interior, 2], vox[interior, 1], vox[interior, 0]].all()
    assert 5000 < occ.sum() < 40000


def test_solid_runs_leaving_the_row(oracle):
    # Known answers from the reference's indexing (cpu_voxelizer.cpp:296-305):
    # a triangle in the plane x = c gives xmax = int(c - 0.5) (unit 1).
    #   c = 9.7 (outside an 8-grid): xmax = 9 -> linear indices base .. base+9,
    #     i.e. the row plus x = 0, 1 of the next row; Morton stops at x = 7.
    #   c = 0.2: q = -0.3 -> int() = 0 -> exactly bit 0 of the row.
    #   c = -0.7: q = -1.2 -> negative run -> empty (counted).
    grid = 8
    info = oracle.make_info([0, 0, 0], [8, 8, 8], grid, 1)
    yz = np.array([[0.2, 0.3], [7.1, 0.6], [0.9, 7.4]], np.float32)
    A, B, C = yz.astype(np.float64)

    def cr(u, v):
        return u[0] * v[1] - u[1] * v[0]

    cols = []
    for z in range(grid):
        for y in range(grid):
            P = np.array([y + 0.5, z + 0.5])
            w = [cr(B - A, P - A), cr(C - B, P - B), cr(A - C, P - C)]
            if all(x > 1e-9 for x in w) or all(x < -1e-9 for x in w):
                cols.append((y, z))
    for c, xmax, counted in ((9.7, 9, 1), (0.2, 0, 0), (-0.7, -1, 1)):
        tri = np.concatenate([np.full((3, 1), c, np.float32), yz], 1)[None]
        for morton in (False, True):
            table, offrow, _ = oracle.solid(info, tri, morton=morton)
            assert offrow == counted * len(cols)
            exp = np.zeros(grid ** 3, bool)
            for (y, z) in cols:
                base = y * grid + z * grid * grid
                if morton:
                    for x in range(0, min(xmax, grid - 1) + 1):
                        exp[oracle.morton_encode(x, y, z)] ^= True
                else:
                    exp[base: base + xmax + 1] ^= True
            got = np.unpackbits(table.astype(">u4").view(np.uint8)).astype(bool)[: grid ** 3]
            assert (got == exp).all(), (c, morton)
