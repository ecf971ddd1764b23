"""CPU check of the conservative x band used to prune the surface raster
(raster.h: plane_band / band_chunks): every voxel whose binary32 plane test
passes must lie inside the band computed for its row.  The float32 plane test
is evaluated here with numpy in the reference's operation order
(cpu_voxelizer.cpp:72-87, 138-140)."""
import numpy as np

F = np.float32
EPS = 5.9604644775390625e-08


def _setup(v, unit):
    e0, e1 = v[1] - v[0], v[2] - v[1]
    cr = np.array([e0[1] * e1[2] - e0[2] * e1[1], e0[2] * e1[0] - e0[0] * e1[2], e0[0] * e1[1] - e0[1] * e1[0]], F)
    s = F(F(cr[0] * cr[0]) + F(cr[1] * cr[1])) + F(cr[2] * cr[2])
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = F(1.0) / np.sqrt(s, dtype=F)
        n = (cr * inv).astype(F)
    c = np.where(n > 0, unit, F(0)).astype(F)
    d1 = F(F(n[0] * F(c[0] - v[0][0])) + F(n[1] * F(c[1] - v[0][1]))) + F(n[2] * F(c[2] - v[0][2]))
    d2 = F(F(n[0] * F(F(unit[0] - c[0]) - v[0][0])) + F(n[1] * F(F(unit[1] - c[1]) - v[0][1]))) + \
        F(n[2] * F(F(unit[2] - c[2]) - v[0][2]))
    return n, d1, d2


def _band(n, d1, d2, unit, xa, xb, yb, zb, py, pz):
    ax = abs(float(n[0]))
    sx = float(F(F(F(abs(n[0])) * F(xb + 1)) * unit[0]) * F(1.0001))
    s = float(n[0]) * float(unit[0])
    C = float(n[1]) * float(py) + float(n[2]) * float(pz)
    dmax, dmin = max(float(d1), float(d2)), min(float(d1), float(d2))
    S = sx + abs(float(n[1]) * float(py)) + abs(float(n[2]) * float(pz)) + abs(dmax) + abs(dmin)
    M = 64.0 * EPS * S + 1e-30
    a, b = sorted(((-dmax - M - C) / s, (-dmin + M - C) / s))
    lo, hi = xa, xb
    if a > xb:
        return xb + 1, xb
    if b < xa:
        return xa, xa - 1
    if a > xa:
        lo = int(np.floor(a)) - 2
    if b < xb:
        hi = int(np.ceil(b)) + 2
    return max(lo, xa), min(hi, xb)


def test_band_contains_every_passing_voxel():
    rng = np.random.default_rng(0)
    checked = 0
    for trial in range(300):
        G = int(rng.choice([64, 256, 1024]))
        unit = np.full(3, F(2.0) / F(G), F) * F(1 + 1e-7 * rng.integers(0, 3))
        c = rng.uniform(0.2, 1.8, 3)
        scale = 10 ** rng.uniform(-3, -0.3)
        v = (c + rng.normal(size=(3, 3)) * scale).astype(F)
        v = np.clip(v, 0.01, 1.99).astype(F)
        n, d1, d2 = _setup(v, unit)
        if not np.all(np.isfinite(n)) or n[0] == 0:
            continue
        lo3 = np.clip((v.min(0) / unit).astype(np.int64), 0, G - 1)
        hi3 = np.clip((v.max(0) / unit).astype(np.int64), 0, G - 1)
        xs = np.arange(lo3[0], hi3[0] + 1)
        px = (xs.astype(F) * unit[0]).astype(F)
        for y in range(lo3[1], hi3[1] + 1, max(1, (hi3[1] - lo3[1]) // 6 + 1)):
            for z in range(lo3[2], hi3[2] + 1, max(1, (hi3[2] - lo3[2]) // 6 + 1)):
                py, pz = F(F(y) * unit[1]), F(F(z) * unit[2])
                t = ((n[0] * px).astype(F) + F(n[1] * py)).astype(F)
                t = (t + F(n[2] * pz)).astype(F)
                r = ((t + d1).astype(F) * (t + d2).astype(F)).astype(F)
                passing = xs[~(r > 0)]
                lo, hi = _band(n, d1, d2, unit, int(lo3[0]), int(hi3[0]), int(hi3[1]), int(hi3[2]), py, pz)
                if passing.size:
                    assert passing.min() >= lo and passing.max() <= hi, (trial, y, z)
                    checked += passing.size
    assert checked > 1000
