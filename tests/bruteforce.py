"""Independent float64 formulations of what the voxelizer computes.

These are NOT restatements of the reference code: they use the textbook
separating-axis theorem (13 axes, two-sided) for triangle/box overlap and
+x ray parity with the reference's truncation rule for the solid fill.  They
pin the oracle away from ties: every voxel whose decision is within ``tol``
voxels of a boundary is reported as a tie and excluded from comparisons.
"""
from __future__ import annotations

import numpy as np


def surface_sat(tris_translated: np.ndarray, unit: np.ndarray, grid: int, tol: float = 1e-4):
    """tris_translated: float64 [N,3,3] (vertex - bbox.min, computed in fp32
    by the caller as the reference does).  unit: float64[3].

    Returns (occupancy bool [z,y,x], tie bool [z,y,x]).

    Voxel (x,y,z) is set iff it lies in the triangle's voxel range
    trunc(min/u) <= x <= trunc(max/u) (clamped) and the closed box
    [x*u, (x+1)*u] overlaps the triangle on the plane axis and the nine
    edge x box-axis axes."""
    G = grid
    occ = np.zeros((G, G, G), dtype=bool)
    tie = np.zeros((G, G, G), dtype=bool)
    u = unit.astype(np.float64)
    axes_box = np.eye(3)
    for tri in tris_translated.astype(np.float64):
        v0, v1, v2 = tri
        mn, mx = tri.min(0), tri.max(0)
        lo = np.clip(np.trunc(mn / u).astype(int), 0, G - 1)
        hi = np.clip(np.trunc(mx / u).astype(int), 0, G - 1)
        xs = np.arange(lo[0], hi[0] + 1)
        ys = np.arange(lo[1], hi[1] + 1)
        zs = np.arange(lo[2], hi[2] + 1)
        Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
        p0 = np.stack([X * u[0], Y * u[1], Z * u[2]], -1).reshape(-1, 3)  # min corners
        # 8 corners per voxel
        offs = np.array([[i, j, k] for i in (0, 1) for j in (0, 1) for k in (0, 1)], dtype=np.float64) * u
        corners = p0[:, None, :] + offs[None, :, :]  # [M,8,3]
        range_tie = False
        for k in range(3):
            for val in (mn[k] / u[k], mx[k] / u[k]):
                if abs(val - round(val)) < tol:
                    range_tie = True
        e = [v1 - v0, v2 - v1, v0 - v2]
        n = np.cross(e[0], e[1])
        axes = []
        if np.linalg.norm(n) > 0:
            axes.append(n)
        for ei in e:
            for b in axes_box:
                a = np.cross(ei, b)
                if np.linalg.norm(a) > 1e-300:
                    axes.append(a)
        # largest separation over all axes: > 0 means disjoint, <= 0 overlap
        maxsep = np.full(p0.shape[0], -np.inf)
        for a in axes:
            a = a / np.linalg.norm(a)
            tp = tri @ a
            tmin, tmax = tp.min(), tp.max()
            cp = corners @ a  # [M,8]
            bmin, bmax = cp.min(1), cp.max(1)
            sep = np.maximum(bmin - tmax, tmin - bmax) / u.max()
            maxsep = np.maximum(maxsep, sep)
        sel = (Z.ravel(), Y.ravel(), X.ravel())
        if range_tie:
            tie[sel] = True
            continue
        occ[sel] |= maxsep < -tol
        tie[sel] |= np.abs(maxsep) <= tol
    tie &= ~occ
    return occ, tie


def solid_parity(tris_translated: np.ndarray, unit: np.ndarray, grid: int, tol: float = 1e-4):
    """Ray-parity fill with the reference's truncation rule.

    For each column centre (yc, zc) = ((y+0.5)u, (z+0.5)u) and each triangle
    whose YZ projection strictly contains it, the plane's x-intercept xc
    toggles voxels x = 0 .. trunc(xc/u - 0.5).  Columns touching a projected
    edge/vertex within tol, or whose xc/u - 0.5 is within tol of an integer,
    or hit by a triangle with |n_x| near the 1e-6 skip threshold, are ties.

    Returns (occupancy bool [z,y,x], tie_column bool [z,y])."""
    G = grid
    u = unit.astype(np.float64)
    par = np.zeros((G, G, G), dtype=bool)
    tiecol = np.zeros((G, G), dtype=bool)
    yc = (np.arange(G) + 0.5) * u[1]
    zc = (np.arange(G) + 0.5) * u[2]
    for tri in tris_translated.astype(np.float64):
        v0, v1, v2 = tri
        n = np.cross(v1 - v0, v2 - v1)
        ln = np.linalg.norm(n)
        if ln == 0:
            continue
        nx = n[0] / ln
        a, b, c = tri[:, 1:]
        ymin, zmin = np.min(tri[:, 1:], 0)
        ymax, zmax = np.max(tri[:, 1:], 0)
        ys = np.nonzero((yc >= ymin - tol * u[1]) & (yc <= ymax + tol * u[1]))[0]
        zs = np.nonzero((zc >= zmin - tol * u[2]) & (zc <= zmax + tol * u[2]))[0]
        if ys.size == 0 or zs.size == 0:
            continue
        Z, Y = np.meshgrid(zs, ys, indexing="ij")
        P = np.stack([yc[Y], zc[Z]], -1)

        def edge(p, q):
            d = q - p
            ln2 = np.hypot(d[0], d[1])
            return ((P[..., 0] - p[0]) * d[1] - (P[..., 1] - p[1]) * d[0]) / max(ln2, 1e-300)

        w0, w1, w2 = edge(a, b), edge(b, c), edge(c, a)
        inside = ((w0 > 0) & (w1 > 0) & (w2 > 0)) | ((w0 < 0) & (w1 < 0) & (w2 < 0))
        near = (np.abs(w0) <= tol * u[1]) | (np.abs(w1) <= tol * u[1]) | (np.abs(w2) <= tol * u[1])
        if abs(abs(nx) - 1e-6) < 1e-7:
            tiecol[Z[inside | near], Y[inside | near]] = True
            continue
        if abs(nx) < 1e-6:
            continue
        tiecol[Z[near], Y[near]] = True
        zi, yi = Z[inside & ~near], Y[inside & ~near]
        if zi.size == 0:
            continue
        # plane n.(p - v0) = 0 -> x = v0.x - (n_y (y - v0.y) + n_z (z - v0.z)) / n_x
        xc = v0[0] - (n[1] * (yc[yi] - v0[1]) + n[2] * (zc[zi] - v0[2])) / n[0]
        q = xc / u[0] - 0.5
        qt = np.trunc(q)
        bad = np.abs(q - np.round(q)) <= tol
        tiecol[zi[bad], yi[bad]] = True
        for zz, yy, xm in zip(zi, yi, qt.astype(int)):
            if xm >= 0:
                par[zz, yy, : min(xm, G - 1) + 1] ^= True
    return par, tiecol


def morton_index(x, y, z):
    """Bit interleave x0 y0 z0 x1 y1 z1 ... of non-negative integer arrays (16 bits each)."""
    x, y, z = (np.asarray(a, np.uint64) for a in (x, y, z))
    out = np.zeros(np.broadcast(x, y, z).shape, np.uint64)
    for b in range(16):
        ub = np.uint64(b)
        out |= ((x >> ub) & np.uint64(1)) << np.uint64(3 * b)
        out |= ((y >> ub) & np.uint64(1)) << np.uint64(3 * b + 1)
        out |= ((z >> ub) & np.uint64(1)) << np.uint64(3 * b + 2)
    return out.astype(np.int64)
