"""Pins of the oracle's restatement of the reference writers (util_io.cpp).

The writers' bodies are checked against what the formats fix independently of
the restatement: the binvox body decodes (run lengths 1..255, each pair
(value, count)) back to the table in the published binvox voxel order -- y
fastest, then z, then x -- and runs are maximal up to the 255 cap; the OBJ
point cloud parses back to exactly the set voxels in x, y, z order with
coordinates v + 0.5; every OBJ cube is 8 corner vertices of a unit cube plus
the 12 fixed faces.  Morton tables of the same voxels give the same bodies."""
import numpy as np
import pytest


def unpack(table: np.ndarray, G: int, morton: bool = False) -> np.ndarray:
    """vox[z, y, x] booleans from a bit table (MSB-first words)."""
    bits = np.unpackbits(table.view(np.uint32).byteswap().view(np.uint8))[: G ** 3].astype(bool)
    if not morton:
        return bits.reshape(G, G, G)
    from tests.bruteforce import morton_index
    out = np.zeros((G, G, G), bool)
    z, y, x = np.meshgrid(np.arange(G), np.arange(G), np.arange(G), indexing="ij")
    out[z, y, x] = bits[morton_index(x, y, z)]
    return out


def pack(vox: np.ndarray, morton: bool = False) -> np.ndarray:
    G = vox.shape[0]
    n = G ** 3
    bits = np.zeros(((n + 31) // 32) * 32, np.uint8)
    if morton:
        from tests.bruteforce import morton_index
        z, y, x = np.nonzero(vox)
        bits[morton_index(x, y, z)] = 1
    else:
        bits[:n] = vox.reshape(-1)
    return np.packbits(bits).view(np.uint32).byteswap()


def random_vox(G: int, seed: int, p: float) -> np.ndarray:
    rng = np.random.default_rng(seed)
    vox = rng.random((G, G, G)) < p
    # long runs too: a solid slab and an empty slab
    vox[: G // 4] = True
    vox[G // 2: G // 2 + 2] = False
    return vox


def decode_binvox(body: bytes, G: int) -> np.ndarray:
    b = np.frombuffer(body, np.uint8)
    assert b.size % 2 == 0
    vals, counts = b[0::2], b[1::2]
    assert counts.min() >= 1 and set(np.unique(vals)) <= {0, 1}
    # maximal runs: a pair may repeat the previous value only after a full 255
    same = vals[1:] == vals[:-1]
    assert (counts[:-1][same] == 255).all()
    flat = np.repeat(vals, counts)
    assert flat.size == G ** 3
    return flat.reshape(G, G, G).astype(bool)  # [x, z, y]


@pytest.mark.parametrize("G,p", [(1, 1.0), (2, 0.5), (5, 0.3), (8, 0.02), (16, 0.5), (24, 0.1)])
def test_binvox_body_decodes_to_table(oracle, G, p):
    vox = random_vox(G, G, p)
    body = oracle.binvox_body(pack(vox), G)
    xzy = decode_binvox(body, G)
    assert (xzy.transpose(1, 2, 0) == vox).all()  # [x, z, y] -> [z, y, x]


def test_binvox_runs_cap_at_255(oracle):
    G = 16  # 4096 voxels, all set: 16 runs of 255 + 16
    body = oracle.binvox_body(pack(np.ones((G, G, G), bool)), G)
    b = np.frombuffer(body, np.uint8)
    assert list(b[1::2]) == [255] * 16 + [16] and set(b[0::2]) == {1}
    body = oracle.binvox_body(pack(np.zeros((G, G, G), bool)), G)
    assert list(np.frombuffer(body, np.uint8)[1::2]) == [255] * 16 + [16]


def test_obj_points_parse_back(oracle):
    G = 12
    vox = random_vox(G, 3, 0.05)
    vox[: G // 4] = False
    vox[0, 1, 2] = True  # z=0, y=1, x=2
    text = oracle.obj_points(pack(vox), G).decode()
    lines = text.splitlines()
    assert all(l.startswith("v ") for l in lines)
    pts = np.array([[float(t) for t in l.split()[1:]] for l in lines])
    z, y, x = np.nonzero(vox)
    order = np.lexsort((z, y, x))  # x slowest, z fastest
    exp = np.stack([x[order], y[order], z[order]], 1) + 0.5
    np.testing.assert_array_equal(pts, exp)
    assert "v 2.5 1.5 0.5" in lines


def test_obj_cubes_structure(oracle):
    G = 6
    vox = np.zeros((G, G, G), bool)
    vox[3, 2, 1] = True
    vox[0, 5, 5] = True
    text = oracle.obj_cubes(pack(vox), G).decode().splitlines()
    assert len(text) == 40
    faces = [l for l in text if l.startswith("f ")]
    assert faces[:12] == faces[12:]
    for blk in (text[:20], text[20:]):
        verts = np.array([[int(t) for t in l.split()[1:]] for l in blk[:8]])
        lo = verts.min(0)
        assert (verts.max(0) - lo == 1).all() and len({tuple(v) for v in verts}) == 8
        # faces index the block's own vertices (relative -1..-8) and cover the 6 sides twice
        idx = np.array([[int(t) for t in l.split()[1:]] for l in blk[8:]])
        assert idx.min() == -8 and idx.max() == -1
    first = np.array([[int(t) for t in l.split()[1:]] for l in text[:8]]).min(0)
    assert tuple(first) == (1, 2, 3)  # x-major order: the voxel at x=1 is written first


@pytest.mark.parametrize("fmt", ["binvox_body", "obj_points", "obj_cubes"])
def test_morton_tables_give_same_bodies(oracle, fmt):
    G = 16
    vox = random_vox(G, 7, 0.03)
    vox[: G // 4] = False
    fn = getattr(oracle, fmt)
    assert fn(pack(vox, morton=True), G, morton=True) == fn(pack(vox), G)


def parse_vox(data: bytes):
    """Independent reader of the MagicaVoxel chunk format: (chunks list, size dict per model)."""
    import struct
    assert data[:4] == b"VOX " and struct.unpack_from("<i", data, 4)[0] == 150
    assert data[8:12] == b"MAIN"
    content, child = struct.unpack_from("<ii", data, 12)
    assert content == 0 and child == len(data) - 20
    pos, chunks = 20, []
    while pos < len(data):
        cid = data[pos:pos + 4].decode()
        n, m = struct.unpack_from("<ii", data, pos + 4)
        assert m == 0
        chunks.append((cid, data[pos + 12: pos + 12 + n]))
        pos += 12 + n
    assert pos == len(data)
    return chunks


def read_dict(b: bytes, o: int):
    import struct
    (n,) = struct.unpack_from("<i", b, o)
    o += 4
    d = {}
    for _ in range(n):
        (lk,) = struct.unpack_from("<i", b, o)
        k = b[o + 4: o + 4 + lk].decode()
        o += 4 + lk
        (lv,) = struct.unpack_from("<i", b, o)
        d[k] = b[o + 4: o + 4 + lv].decode()
        o += 4 + lv
    return d, o


@pytest.mark.parametrize("G,p", [(8, 0.05), (130, 0.0002), (260, 0.00002)])
def test_vox_file_structure(oracle, G, p):
    """The .vox voxels are the set voxels fed as (x, G - z, y) and split into
    126^3 cubes; cube order follows the first voxel fed; the scene graph links
    one transform + shape per cube."""
    import struct
    rng = np.random.default_rng(G)
    vox = rng.random((G, G, G)) < p
    vox[0, 0, 0] = vox[G - 1, G - 1, G - 1] = True
    data = oracle.vox_file(pack(vox), G)
    chunks = parse_vox(data)
    names = [c[0] for c in chunks]
    ncubes = names.count("XYZI")
    assert names[: 2 * ncubes] == ["SIZE", "XYZI"] * ncubes
    assert names[2 * ncubes:] == ["nTRN", "nGRP"] + ["nTRN", "nSHP"] * ncubes + ["RGBA"]
    # voxels of every cube, reconstructed to grid coordinates via the translations' cube grid
    z, y, x = np.nonzero(vox)
    order = np.lexsort((z, y, x))
    X, Y, Z = x[order], G - z[order], y[order]
    cube = (X // 126) * 10000 + (Y // 126) * 100 + Z // 126
    _, first = np.unique(cube, return_index=True)
    cube_order = cube[np.sort(first)]
    for k, key in enumerate(cube_order):
        assert struct.unpack_from("<iii", chunks[2 * k][1]) == (126, 126, 126)
        body = chunks[2 * k + 1][1]
        (n,) = struct.unpack_from("<i", body)
        got = np.frombuffer(body[4:], np.uint8).reshape(n, 4)
        sel = cube == key
        exp = np.stack([X[sel] % 126, Y[sel] % 126, Z[sel] % 126, np.ones(sel.sum(), int)], 1)
        assert (got == exp).all()
    grp = chunks[2 * ncubes + 1][1]
    node, = struct.unpack_from("<i", grp)
    _, o = read_dict(grp, 4)
    (cnt,) = struct.unpack_from("<i", grp, o)
    kids = struct.unpack_from(f"<{cnt}i", grp, o + 4)
    assert node == 1 and cnt == ncubes and kids == tuple(range(2, 2 + 2 * ncubes, 2))
    for k in range(ncubes):
        trn = chunks[2 * ncubes + 2 + 2 * k][1]
        nid, = struct.unpack_from("<i", trn)
        _, o = read_dict(trn, 4)
        child, res, layer, nfr = struct.unpack_from("<iiii", trn, o)
        frame, _ = read_dict(trn, o + 16)
        assert (nid, child, res, layer, nfr) == (2 + 2 * k, 3 + 2 * k, -1, 0, 1)
        assert len(frame["_t"].split()) == 3
        shp = chunks[2 * ncubes + 3 + 2 * k][1]
        nid, = struct.unpack_from("<i", shp)
        _, o = read_dict(shp, 4)
        nmod, mid = struct.unpack_from("<ii", shp, o)
        attrs, _ = read_dict(shp, o + 8)
        assert (nid, nmod, mid, attrs) == (3 + 2 * k, 1, k, {"_f": "0"})
    pal = np.frombuffer(chunks[-1][1], np.uint32)
    assert pal[0] == 0x00FFFFFF and not pal[1:].any()


def test_vox_translation_centres_single_cube(oracle):
    """One cube: _t = floor((0.5 * 126) - lo - size/2) for x and y, floor(63) for z."""
    G = 8
    vox = np.zeros((G, G, G), bool)
    vox[2, 3, 1] = vox[5, 6, 4] = True  # (z, y, x)
    chunks = parse_vox(oracle.vox_file(pack(vox), G))
    trn = chunks[4][1]
    _, o = read_dict(trn, 4)
    frame, _ = read_dict(trn, o + 16)
    X = [1, 4]; Y = [G - 2, G - 5]
    tx = int(np.floor(63.0 - min(X) - (max(X) - min(X)) * 0.5))
    ty = int(np.floor(63.0 - min(Y) - (max(Y) - min(Y)) * 0.5))
    assert frame["_t"] == f"{tx} {ty} 63"


def test_vox_empty_table(oracle):
    chunks = parse_vox(oracle.vox_file(pack(np.zeros((4, 4, 4), bool)), 4))
    assert [c[0] for c in chunks] == ["nTRN", "nGRP", "RGBA"]
