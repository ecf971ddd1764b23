"""Deterministic synthetic meshes for the voxelization workloads.

Every generator returns float32 data and is a pure function of its seed, so
the GPU path and the CPU oracle can be fed the same bytes in one process.

Workloads (all triangles FP32, AoS ``[N, 3, 3]`` = v0.xyz, v1.xyz, v2.xyz, the
layout the reference packs for its kernels, main.cpp:61-80):

* ``icosphere_jitter(level=4)``       - 64^3 surface workload: icosphere with
  2,562 vertices / 5,120 triangles, unit radius, every vertex scaled by a
  seeded factor 1 + U(-0.02, 0.02).
* ``torus_slivers()``                 - 512^3 Morton surface workload: torus
  R=1, r=0.35, 1024x512 quads (1,048,576 triangles), per-coordinate Gaussian
  jitter sigma = 0.001 r, and 10% of the triangles turned into slivers (one
  edge collapsed to 1e-4 of the voxel size).
* ``icosphere_sh(level=8)``           - 1024^3 solid workload: watertight
  icosphere (1,310,720 triangles) displaced radially by a sum of 8 real
  spherical harmonics of degree <= 6, sum |a_k| <= 0.15.
* ``batch_meshes(256)``               - 256 meshes with log-uniform triangle
  counts in [500, 500k], grid sizes from {32, 64, 128, 256}, an equal mix of
  jittered icospheres, tori and random triangle soups.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

__all__ = [
    "icosphere", "icosphere_jitter", "icosphere_sh", "torus", "torus_slivers",
    "triangle_soup", "to_aos", "batch_meshes", "BatchMesh", "read_obj",
]


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def icosphere(level: int) -> tuple[np.ndarray, np.ndarray]:
    """Unit icosphere: icosahedron + ``level`` midpoint subdivisions.

    Returns (vertices float64 [V,3], faces int64 [F,3]); V = 10*4^L + 2,
    F = 20*4^L.  Faces are consistently wound (outward normals).
    """
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    for _ in range(level):
        nv = v.shape[0]
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        edges = np.stack([np.stack([a, b], 1), np.stack([b, c], 1), np.stack([c, a], 1)], 0)  # [3,F,2]
        lo = np.minimum(edges[..., 0], edges[..., 1])
        hi = np.maximum(edges[..., 0], edges[..., 1])
        key = lo * nv + hi
        uniq, inv = np.unique(key.ravel(), return_inverse=True)
        ulo, uhi = uniq // nv, uniq % nv
        mid = v[ulo] + v[uhi]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        mids = (inv.reshape(3, -1) + nv)
        ab, bc, ca = mids[0], mids[1], mids[2]
        v = np.concatenate([v, mid], 0)
        f = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1),
                            np.stack([c, ca, bc], 1), np.stack([ab, bc, ca], 1)], 0)
    return v, f


def to_aos(vertices: np.ndarray, faces: np.ndarray) -> np.ndarray:
    """Indexed mesh -> AoS float32 [F,3,3] (main.cpp:61-80 packing)."""
    return np.ascontiguousarray(vertices.astype(np.float32)[faces])


def icosphere_jitter(level: int = 4, amp: float = 0.02, seed: int = 0xC1) -> tuple[np.ndarray, np.ndarray]:
    """Icosphere with every vertex scaled radially by 1 + U(-amp, amp)."""
    v, f = icosphere(level)
    r = 1.0 + _rng(seed).uniform(-amp, amp, size=(v.shape[0], 1))
    return (v * r).astype(np.float32), f.astype(np.int32)


def _real_sph_harm(l: int, m: int, x, y, z):
    """Real spherical harmonic (unnormalised) of degree l, order m, evaluated
    on unit vectors via the associated-Legendre recurrence in cos(theta)=z and
    cos/sin(m phi) built from (x + iy)^m -- only polynomials, no trig calls."""
    am = abs(m)
    # P_m^m(z) up to a constant: (1 - z^2)^{m/2} is absorbed by using (x+iy)^m
    pmm = np.ones_like(z)
    if l == am:
        leg = pmm
    else:
        pm1 = (2 * am + 1) * z * pmm
        if l == am + 1:
            leg = pm1
        else:
            p0, p1 = pmm, pm1
            for ll in range(am + 2, l + 1):
                p0, p1 = p1, ((2 * ll - 1) * z * p1 - (ll + am - 1) * p0) / (ll - am)
            leg = p1
    cz = (x + 1j * y) ** am
    ang = cz.real if m >= 0 else cz.imag
    return leg * ang


def icosphere_sh(level: int = 8, n_terms: int = 8, max_degree: int = 6, amp_total: float = 0.15,
                 seed: int = 0xC3) -> tuple[np.ndarray, np.ndarray]:
    """Watertight icosphere displaced radially by a seeded low-frequency field.

    r(dir) = 1 + sum_k a_k * Y_k(dir) / max|Y_k|, with sum |a_k| = amp_total,
    so r stays in [1 - amp_total, 1 + amp_total] (> 0.5)."""
    v, f = icosphere(level)
    rng = _rng(seed)
    x, y, z = v[:, 0], v[:, 1], v[:, 2]
    w = rng.uniform(0.2, 1.0, size=n_terms)
    a = w / w.sum() * amp_total * rng.choice([-1.0, 1.0], size=n_terms)
    r = np.ones_like(x)
    for k in range(n_terms):
        l = int(rng.integers(1, max_degree + 1))
        m = int(rng.integers(-l, l + 1))
        yk = _real_sph_harm(l, m, x, y, z)
        mx = np.max(np.abs(yk))
        if mx > 0:
            r += a[k] * yk / mx
    return (v * r[:, None]).astype(np.float32), f.astype(np.int32)


def torus(nu: int, nv: int, R: float = 1.0, r: float = 0.35) -> tuple[np.ndarray, np.ndarray]:
    """Torus with nu (major) x nv (minor) quads, two triangles each."""
    u = np.arange(nu, dtype=np.float64) * (2.0 * math.pi / nu)
    w = np.arange(nv, dtype=np.float64) * (2.0 * math.pi / nv)
    cu, su = np.cos(u)[:, None], np.sin(u)[:, None]
    cw, sw = np.cos(w)[None, :], np.sin(w)[None, :]
    rad = R + r * cw
    px, py = rad * cu, rad * su
    verts = np.stack([px, py, np.broadcast_to(r * sw, px.shape)], -1).reshape(-1, 3)
    i = np.arange(nu)[:, None]
    j = np.arange(nv)[None, :]
    a = i * nv + j
    b = ((i + 1) % nu) * nv + j
    c = ((i + 1) % nu) * nv + (j + 1) % nv
    d = i * nv + (j + 1) % nv
    faces = np.concatenate([np.stack([a, b, c], -1).reshape(-1, 3),
                            np.stack([a, c, d], -1).reshape(-1, 3)], 0)
    # interleave so triangles of one quad are adjacent in memory (spatial locality)
    half = faces.shape[0] // 2
    inter = np.empty_like(faces)
    inter[0::2], inter[1::2] = faces[:half], faces[half:]
    return verts, inter


def torus_slivers(nu: int = 1024, nv: int = 512, R: float = 1.0, r: float = 0.35, grid: int = 512,
                  sliver_frac: float = 0.10, seed: int = 0xC2) -> np.ndarray:
    """Jittered torus as AoS triangles with a seeded 10% of them turned into
    slivers: vertex 1 is moved onto vertex 0 plus a random direction of length
    1e-4 voxel, so edge e0 is ~1e-4 voxel long."""
    rng = _rng(seed)
    v, f = torus(nu, nv, R, r)
    v = v + rng.normal(0.0, 0.001 * r, size=v.shape)
    tris = v[f]  # float64 [F,3,3]
    extent = (tris.reshape(-1, 3).max(0) - tris.reshape(-1, 3).min(0)).max()
    unit = extent / grid
    nsl = int(round(sliver_frac * tris.shape[0]))
    idx = rng.choice(tris.shape[0], size=nsl, replace=False)
    d = rng.normal(size=(nsl, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    tris[idx, 1] = tris[idx, 0] + d * (1e-4 * unit)
    return np.ascontiguousarray(tris.astype(np.float32))


def triangle_soup(n: int, seed: int, mu: float = -3.0, sigma: float = 1.0) -> np.ndarray:
    """Random soup: v0 ~ U[0,1]^3, v1/v2 = v0 + lognormal(mu, sigma) * random unit dir."""
    rng = _rng(seed)
    v0 = rng.uniform(0.0, 1.0, size=(n, 3))
    out = np.empty((n, 3, 3), dtype=np.float64)
    out[:, 0] = v0
    for k in (1, 2):
        d = rng.normal(size=(n, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        out[:, k] = v0 + d * rng.lognormal(mu, sigma, size=(n, 1))
    return np.ascontiguousarray(out.astype(np.float32))


@dataclass
class BatchMesh:
    tris: np.ndarray      # float32 [N,3,3]
    grid: int             # cubic grid length
    solid: bool           # closed meshes are voxelized solid, soups surface
    kind: str


def batch_meshes(count: int = 256, seed: int = 0xC4, n_min: int = 500, n_max: int = 500_000,
                 grids=(32, 64, 128, 256)) -> list[BatchMesh]:
    """Batched workload: count meshes, N ~ round(exp(U(ln n_min, ln n_max))),
    grid uniform over ``grids``, type by index mod 3 (icosphere, torus, soup)."""
    rng = _rng(seed)
    out: list[BatchMesh] = []
    for i in range(count):
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        g = int(rng.choice(grids))
        kind = ("icosphere", "torus", "soup")[i % 3]
        sub = int(rng.integers(0, 2**31))
        if kind == "icosphere":
            level = int(np.clip(round(math.log(max(n, 20) / 20.0, 4)), 0, 7))
            v, f = icosphere_jitter(level=level, amp=0.02, seed=sub)
            out.append(BatchMesh(to_aos(v, f), g, True, kind))
        elif kind == "torus":
            a = max(3, int(round(math.sqrt(n))))
            b = max(3, int(round(n / (2 * a))))
            v, f = torus(a, b)
            v = v + _rng(sub).normal(0.0, 1e-4, size=v.shape)
            out.append(BatchMesh(to_aos(v, f), g, True, kind))
        else:
            out.append(BatchMesh(triangle_soup(n, sub), g, False, kind))
    return out


def read_obj(path: str) -> tuple[np.ndarray, np.ndarray]:
    """Minimal Wavefront OBJ reader: ``v x y z`` and triangle/polygon ``f``
    records (``i``, ``i/t``, ``i//n``, ``i/t/n``); polygons are fanned.
    Returns (vertices float32 [V,3], faces int32 [F,3])."""
    verts: list[tuple[float, float, float]] = []
    faces: list[tuple[int, int, int]] = []
    with open(path, "r") as fh:
        for line in fh:
            if line.startswith("v "):
                p = line.split()
                verts.append((float(p[1]), float(p[2]), float(p[3])))
            elif line.startswith("f "):
                idx = []
                for tok in line.split()[1:]:
                    k = int(tok.split("/")[0])
                    idx.append(k - 1 if k > 0 else len(verts) + k)
                for j in range(1, len(idx) - 1):
                    faces.append((idx[0], idx[j], idx[j + 1]))
    return np.asarray(verts, dtype=np.float32), np.asarray(faces, dtype=np.int32)
