"""Hand-derived known answers at the solid voxelizer's thresholds.

The reference's solid column test (cpu_voxelizer.cpp:196-238 and :265,
voxelize_solid.cu:13-92) makes three decisions that only inputs placed
exactly on them pin:
  * the skip of a triangle with (double)|n.x| < 1e-6, n = normalize(cross(e0, e1));
  * check_point_triangle's on-edge test, (double)|t| < 1e-6 together with the
    edge's bounding-box products <= 0, which hands the column to the edge's
    top-left rule instead of the strict inside test;
  * TopLeftEdge(v0, v1) = v1.y < v0.y || (v1.y == v0.y && v0.x > v1.x) on the
    counter-clockwise YZ triangle, whose equal-y branch decides horizontal edges.

Every case is one triangle on a 32^3 grid over [0, 32]^3 (unit 1, column
(y, z) has its centre at (y + 0.5, z + 0.5)).  Each coordinate is a float32
value, and the float arithmetic the reference performs on these inputs is
exact (each case says why), so the expected tables follow from the numbers
written here.  An expected table is given as {(y, z): xmax}: column (y, z)
holds bits x = 0 .. xmax, all other bits are clear.

The "float" below is IEEE binary32: 0x1.0c6f7ap-20 is the float just below
1e-6 (the value of 1e-6f) and 0x1.0c6f7cp-20 the float just above it.
"""
from dataclasses import dataclass
from fractions import Fraction

import numpy as np

GRID = 32
BOX_MIN = (0.0, 0.0, 0.0)
BOX_MAX = (32.0, 32.0, 32.0)

F_BELOW = float.fromhex("0x1.0c6f7ap-20")   # 1e-6f, below 1e-6
F_ABOVE = float.fromhex("0x1.0c6f7cp-20")   # next float, above 1e-6
X_PLANE = float(np.float32(5.3))            # the plane x = 5.3f: get_x_coordinate = 5.3f, xmax = 4


@dataclass
class Kat:
    name: str
    tri: np.ndarray   # (1, 3, 3) float32, vertex order as given to the voxelizer
    cols: dict        # {(y, z): xmax}

    def occupancy(self) -> np.ndarray:
        """Expected occupancy, indexed [z, y, x]."""
        occ = np.zeros((GRID, GRID, GRID), bool)
        for (y, z), xmax in self.cols.items():
            occ[z, y, : xmax + 1] = True
        return occ


def _tri(vertices) -> np.ndarray:
    t = np.asarray(vertices, np.float64)
    f = t.astype(np.float32)
    assert (f.astype(np.float64) == t).all(), "every coordinate must be a float32 value"
    return f.reshape(1, 3, 3)


def _orders(v):
    """The six vertex orders of one triangle: three rotations of each winding."""
    a, b, c = v
    return [("abc", (a, b, c)), ("bca", (b, c, a)), ("cab", (c, a, b)),
            ("acb", (a, c, b)), ("cba", (c, b, a)), ("bac", (b, a, c))]


def normal_threshold_cases():
    """|n.x| exactly at the float below and the float above 1e-6.

    v0 = (x0, 1, 2), v1 = (x0 + 2^20, 1, 2 - h), v2 = (x0 + 2^20, 3, 2 - h) with
    h = c * 2^20, x0 = -499996.  All coordinates are float32 values and
    e0 = v1 - v0 = (2^20, 0, -h), e1 = v2 - v1 = (0, 2, 0) exactly, so
    cross(e0, e1) = (2h, -0, 2^21); its squared length 4h^2 + 2^42 rounds to
    2^42, whose reciprocal square root 2^-21 is exact: n = (c, -0, 1).
      c = 0x1.0c6f7ap-20: (double)c < 1e-6, the triangle is skipped (a float
        compare against 1e-6f would not skip it).
      c = 0x1.0c6f7cp-20: the triangle is voxelized.  Its YZ projection
        (1, 2), (1, 2 - h), (3, 2 - h) is counter-clockwise; columns
        y in {1, 2}, z = 1 are tested: (1.5, 1.5) is inside (t1, t2, t3 > 0.27),
        (2.5, 1.5) outside (t3 = 0.25 - 1.5 * 0.548576 < 0).  At (1.5, 1.5)
        get_x_coordinate = -((-0) * 0.5 + 1 * (1.5 - 2)) / c + x0
        = fl(0.5 / c) + x0 = 499999.9375 - 499996 = 3.9375, so
        xmax = int(3.9375 - 0.5) = 3.
    The order (v0, v2, v1) flips n to (-c, -0, -1) and keeps v0, so
    get_x_coordinate = -(0.5) / -c + x0, the same float.
    """
    x0 = -499996.0
    out = []
    for c, cols in ((F_BELOW, {}), (F_ABOVE, {(1, 1): 3})):
        h = c * 2.0 ** 20
        v0 = (x0, 1.0, 2.0)
        v1 = (x0 + 2.0 ** 20, 1.0, 2.0 - h)
        v2 = (x0 + 2.0 ** 20, 3.0, 2.0 - h)
        for tag, order in (("v0v1v2", (v0, v1, v2)), ("v0v2v1", (v0, v2, v1))):
            out.append(Kat(f"nx={c.hex()}-{tag}", _tri(order), dict(cols)))
    return out


def edge_threshold_cases():
    """|t| exactly at the float below and the float above 1e-6.

    In the plane x = 5.3 (so get_x_coordinate = 5.3 and xmax = int(4.8) = 4):
    V0 = (1.5 + 2^-20, 0.25), V1 = (1.5, 0.5 + m * 2^-23) in (y, z), and the
    column centre P = (1.5, 0.5).  P - V0 = (-2^-20, 0.25) and
    P - V1 = (0, -m * 2^-23) exactly, so t for the edge V0 -> V1 is
    (-2^-20) * (-m * 2^-23) - 0.25 * 0 = m * 2^-43 exactly:
      m = 8796093: t = 0x1.0c6f7ap-20 (below 1e-6, on the edge);
      m = 8796094: t = 0x1.0c6f7cp-20 (above 1e-6, not on the edge).
    The bounding-box products are (-2^-20) * 0 <= 0 and 0.25 * (-1.05) <= 0.
    The centre Q = (1.5, 1.5) has |t| = (m * 2^-23 - 1) * 2^-20 = 4.6e-8: on the
    edge for both m.
      side "a": third vertex W = (0.25, 1.0).  (V0, V1, W) is counter-clockwise,
        the edge runs V0 -> V1 upwards (not top-left) and P, Q lie inside it
        (other edge functions >= 0.31).  m = 8796093: P and Q are on an excluded
        edge, the table is empty (a float compare would set P).  m = 8796094:
        P is strictly inside, Q still on the edge: only P.
      side "b": W = (2.75, 1.0).  The triangle is clockwise as given, the
        counter-clockwise order runs the edge V1 -> V0 downwards (top-left), and
        P, Q lie outside it by |t|.  m = 8796093: P and Q are on an included
        edge (a float compare would drop P).  m = 8796094: P is outside, only Q.
    The other centres in the columns the triangles span, (0.5, 0.5), (0.5, 1.5)
    for "a" and (2.5, 0.5), (2.5, 1.5) for "b", are outside by more than 0.1.
    Every vertex order gives the same columns (the edge keeps its direction in
    the counter-clockwise order, and t is computed from the same two vertices).
    """
    out = []
    x = X_PLANE
    for m, cols_a, cols_b in ((8796093, {}, {(1, 0): 4, (1, 1): 4}),
                              (8796094, {(1, 0): 4}, {(1, 1): 4})):
        v0 = (x, 1.5 + 2.0 ** -20, 0.25)
        v1 = (x, 1.5, 0.5 + m * 2.0 ** -23)
        for side, w, cols in (("a", (x, 0.25, 1.0), cols_a), ("b", (x, 2.75, 1.0), cols_b)):
            for tag, order in _orders((v0, v1, w)):
                out.append(Kat(f"t-m{m}-{side}-{tag}", _tri(order), dict(cols)))
    return out


# Triangles with column centres exactly on each of their edges (t = 0 in float:
# every coordinate and difference is a multiple of 0.5 below 8, every product
# and sum exact), vertices off the centres, in the plane x = 5.3 (xmax = 4).
#   T1: A = (0, 0.5), B = (6, 0.5), C = (3, 6.5), counter-clockwise as given.
#     A -> B is horizontal, running +y: the equal-y branch is false, excluded.
#     B -> C (z = 12.5 - 2y) runs upwards: excluded.
#     C -> A (z = 2y + 0.5) runs downwards: included.
#   T2: the mirror z -> 8 - z of T1: A' = (0, 7.5), B' = (6, 7.5), C' = (3, 1.5),
#     clockwise as given; counter-clockwise it is A' -> C' -> B' -> A'.
#     B' -> A' is horizontal, running -y: the equal-y branch is true, included.
#     A' -> C' (z = 7.5 - 2y) runs downwards: included.
#     C' -> B' (z = 2y - 4.5) runs upwards: excluded.
_EDGE_TRIANGLES = {
    "T1": (((0.0, 0.5), (6.0, 0.5), (3.0, 6.5)),
           {"AB": ([(y + 0.5, 0.5) for y in range(6)], False),
            "BC": ([(5.5, 1.5), (4.5, 3.5), (3.5, 5.5)], False),
            "CA": ([(0.5, 1.5), (1.5, 3.5), (2.5, 5.5)], True)}),
    "T2": (((0.0, 7.5), (6.0, 7.5), (3.0, 1.5)),
           {"AB": ([(y + 0.5, 7.5) for y in range(6)], True),
            "BC": ([(3.5, 2.5), (4.5, 4.5), (5.5, 6.5)], False),
            "CA": ([(0.5, 6.5), (1.5, 4.5), (2.5, 2.5)], True)}),
}


def _cross(o, a, b):
    return (a[0] - o[0]) * (b[1] - o[1]) - (a[1] - o[1]) * (b[0] - o[0])


def _on_segment(p, a, b):
    return _cross(a, b, p) == 0 and min(a[0], b[0]) <= p[0] <= max(a[0], b[0]) \
        and min(a[1], b[1]) <= p[1] <= max(a[1], b[1])


def _edge_tie_columns(yz, edges):
    """Expected columns of one edge triangle: the centres strictly inside
    (exact rational edge functions) plus the centres on its included edges,
    after checking that the centres on its edges are exactly the listed ones."""
    pts = [tuple(Fraction(c) for c in p) for p in yz]
    a, b, c = pts
    if _cross(a, b, c) < 0:
        b, c = c, b
    listed = {p for cs, _ in edges.values() for p in cs}
    cols = {}
    for z in range(GRID):
        for y in range(GRID):
            p = (Fraction(2 * y + 1, 2), Fraction(2 * z + 1, 2))
            on = any(_on_segment(p, u, v) for u, v in ((a, b), (b, c), (c, a)))
            assert on == ((float(p[0]), float(p[1])) in listed), (y, z)
            inside = _cross(a, b, p) > 0 and _cross(b, c, p) > 0 and _cross(c, a, p) > 0
            if inside:
                cols[(y, z)] = 4
    for cs, included in edges.values():
        for (py, pz) in cs:
            if included:
                cols[(int(py), int(pz))] = 4
    return cols


def edge_tie_cases():
    """Column centres with t exactly 0 on each of the three edges, including
    horizontal edges decided by the equal-y branch; all six vertex orders, so
    each geometric edge is check 1, 2 and 3 of check_point_triangle in turn."""
    out = []
    x = X_PLANE
    for name, (yz, edges) in _EDGE_TRIANGLES.items():
        cols = _edge_tie_columns(yz, edges)
        verts = [(x, y, z) for (y, z) in yz]
        for tag, order in _orders(verts):
            out.append(Kat(f"{name}-{tag}", _tri(order), dict(cols)))
    return out


def all_cases():
    return normal_threshold_cases() + edge_threshold_cases() + edge_tie_cases()


__all__ = ["GRID", "BOX_MIN", "BOX_MAX", "F_BELOW", "F_ABOVE", "Kat", "all_cases",
           "normal_threshold_cases", "edge_threshold_cases", "edge_tie_cases"]
