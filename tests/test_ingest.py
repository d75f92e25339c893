"""Scene ingestion on the CPU: the C ABI's PNG decoder (host code, no GPU)
against images written here with zlib in every colour type, bit depth, filter
and deflate block kind; the OBJ reader against arrays fixed by hand; the oracle's
mesh bake (STriangleMesh::create, CSTriangleMesh.cpp:11-39) and texel decode
(gammaCorrectInv, CSTextureProcess.h:517-522) against closed forms."""
import struct
import zlib

import numpy as np
import pytest

import pyoracle as po
from woodenpbrengine_amd import lib

# ------------------------------------------------------------------ PNG ----


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    return b if pb <= pc else c


def _filter_rows(rows, bpp, kinds):
    """PNG scanline filters (encoder side) of a list of byte rows."""
    out = bytearray()
    prev = bytes(len(rows[0])) if rows else b""
    for y, row in enumerate(rows):
        ft = kinds[y % len(kinds)]
        f = bytearray(len(row))
        for i, x in enumerate(row):
            a = row[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            pred = [0, a, b, (a + b) >> 1, _paeth(a, b, c)][ft]
            f[i] = (x - pred) & 0xFF
        out.append(ft)
        out += f
        prev = row
    return bytes(out)


def _pack(samples, depth):
    """One row of samples (ints) at a bit depth, most significant bits first."""
    if depth == 16:
        return b"".join(struct.pack(">H", int(s)) for s in samples)
    if depth == 8:
        return bytes(int(s) for s in samples)
    bits = "".join(format(int(s), "0%db" % depth) for s in samples)
    bits += "0" * (-len(bits) % 8)
    return bytes(int(bits[i:i + 8], 2) for i in range(0, len(bits), 8))


def _chunk(t, data):
    return struct.pack(">I", len(data)) + t + data + struct.pack(">I", zlib.crc32(t + data) & 0xFFFFFFFF)


def _png(samples, depth, color, kinds=(0,), interlace=False, palette=None, strategy="default"):
    """samples: H x W x channels ints.  strategy: stored / fixed / default."""
    h, w, ch = samples.shape
    bpp = max(1, ch * depth // 8)
    if not interlace:
        raw = _filter_rows([_pack(samples[y].reshape(-1), depth) for y in range(h)], bpp, kinds)
    else:
        raw = b""
        for x0, y0, dx, dy in [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2),
                               (0, 1, 1, 2)]:
            sub = samples[y0::dy, x0::dx]
            if sub.shape[0] and sub.shape[1]:
                raw += _filter_rows([_pack(sub[y].reshape(-1), depth) for y in range(sub.shape[0])], bpp, kinds)
    if strategy == "stored":
        z = zlib.compress(raw, 0)
    elif strategy == "fixed":
        co = zlib.compressobj(6, zlib.DEFLATED, 15, 9, zlib.Z_FIXED)
        z = co.compress(raw) + co.flush()
    else:
        z = zlib.compress(raw, 9)
    ihdr = struct.pack(">IIBBBBB", w, h, depth, color, 0, 0, 1 if interlace else 0)
    out = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", ihdr)
    if palette is not None:
        out += _chunk(b"PLTE", bytes(np.asarray(palette, np.uint8).reshape(-1)))
    out += _chunk(b"tEXt", b"Comment\x00ancillary chunks are skipped")
    # IDAT split in two chunks: the decoder must concatenate them
    out += _chunk(b"IDAT", z[: len(z) // 2]) + _chunk(b"IDAT", z[len(z) // 2:])
    return out + _chunk(b"IEND", b"")


def _decode(tmp_path, data, name="t.png"):
    p = tmp_path / name
    p.write_bytes(data)
    return lib.read_png(p)


@pytest.mark.parametrize("strategy", ["stored", "fixed", "default"])
@pytest.mark.parametrize("kinds", [(0,), (1,), (2,), (3,), (4,), (0, 1, 2, 3, 4)])
def test_png_rgb8_filters_and_blocks(tmp_path, strategy, kinds):
    rng = np.random.default_rng(7)
    img = rng.integers(0, 256, (23, 37, 3))
    img[5:9] = img[4]  # repeated rows: long back-references
    out = _decode(tmp_path, _png(img, 8, 2, kinds, strategy=strategy))
    assert out.shape == (23, 37, 3)
    assert np.array_equal(out, img.astype(np.uint8))


def test_png_large_back_references(tmp_path):
    # a 256 x 256 tile pattern repeating every 96 rows: distances near 32 KiB
    rng = np.random.default_rng(3)
    tile = rng.integers(0, 256, (96, 256, 3))
    img = np.concatenate([tile, tile, tile[:64]], axis=0)
    out = _decode(tmp_path, _png(img, 8, 2, (0, 2, 4)))
    assert np.array_equal(out, img.astype(np.uint8))


@pytest.mark.parametrize("depth", [1, 2, 4, 8, 16])
def test_png_grey(tmp_path, depth):
    rng = np.random.default_rng(depth)
    g = rng.integers(0, 1 << depth, (9, 13, 1))
    out = _decode(tmp_path, _png(g, depth, 0, (0, 1, 4)))
    if depth == 16:
        want = g[..., 0] >> 8
    else:
        want = g[..., 0] * 255 // ((1 << depth) - 1)
    assert np.array_equal(out, np.repeat(want[..., None], 3, axis=2).astype(np.uint8))


@pytest.mark.parametrize("depth", [1, 2, 4, 8])
def test_png_palette(tmp_path, depth):
    rng = np.random.default_rng(10 + depth)
    npal = 1 << depth
    pal = rng.integers(0, 256, (npal, 3))
    idx = rng.integers(0, npal, (11, 7, 1))
    out = _decode(tmp_path, _png(idx, depth, 3, (0, 3), palette=pal))
    assert np.array_equal(out, pal[idx[..., 0]].astype(np.uint8))


@pytest.mark.parametrize("color,depth", [(4, 8), (4, 16), (2, 16), (6, 8), (6, 16)])
def test_png_alpha_and_16_bit(tmp_path, color, depth):
    ch = {2: 3, 4: 2, 6: 4}[color]
    rng = np.random.default_rng(color * 100 + depth)
    s = rng.integers(0, 1 << depth, (6, 10, ch))
    out = _decode(tmp_path, _png(s, depth, color, (1, 2, 3, 4)))
    hi = s >> 8 if depth == 16 else s
    want = np.repeat(hi[..., :1], 3, axis=2) if color == 4 else hi[..., :3]
    assert np.array_equal(out, want.astype(np.uint8))


@pytest.mark.parametrize("shape", [(1, 1), (2, 3), (11, 13), (16, 16)])
def test_png_adam7(tmp_path, shape):
    rng = np.random.default_rng(shape[0] * 31 + shape[1])
    img = rng.integers(0, 256, shape + (3,))
    out = _decode(tmp_path, _png(img, 8, 2, (0, 1, 2, 3, 4), interlace=True))
    assert np.array_equal(out, img.astype(np.uint8))
    g = rng.integers(0, 4, shape + (1,))
    out = _decode(tmp_path, _png(g, 2, 0, (4,), interlace=True), "g.png")
    assert np.array_equal(out[..., 0], (g[..., 0] * 85).astype(np.uint8))


def test_png_round_trip_with_writer(tmp_path):
    rng = np.random.default_rng(5)
    img = rng.integers(0, 256, (17, 29, 3), dtype=np.uint8)
    lib.write_png(tmp_path / "w.png", img)
    assert np.array_equal(lib.read_png(tmp_path / "w.png"), img)


def test_png_palette_index_beyond_palette_is_black(tmp_path):
    # lodepng fills the 256 palette entries with black before reading PLTE and
    # does not check indices (lodepng.cpp:2607-2620, 3378-3384)
    idx = np.array([[0, 3], [1, 2]]).reshape(2, 2, 1)
    out = _decode(tmp_path, _png(idx, 2, 3, palette=[[1, 2, 3], [40, 50, 60]]))
    exp = np.array([[[1, 2, 3], [0, 0, 0]], [[40, 50, 60], [0, 0, 0]]], np.uint8)
    assert np.array_equal(out, exp)


def test_png_errors(tmp_path):
    good = _png(np.zeros((4, 4, 3), int), 8, 2)
    bad_crc = bytearray(good)
    bad_crc[8 + 8 + 13] ^= 1  # IHDR CRC
    cases = {"crc": bytes(bad_crc), "sig": b"\x00" + good[1:], "trunc": good[:60]}
    for name, data in cases.items():
        with pytest.raises(lib.WpbrError):
            _decode(tmp_path, data, name + ".png")
    with pytest.raises(lib.WpbrError):
        lib.read_png(tmp_path / "missing.png")


# ------------------------------------------------------------------ OBJ ----

OBJ_TEXT = """# a triangle and a quad, then a second object
v 0 0 0
v 1 0 0\r
v 1 1 0.5
v 0 1 -2
vt 0 0
vt 1 0
vt 1 0.25
vt 0 1
vn 0 0 1
vn 0 1 0
f 1/1/1 2/2/1 3/3/2
f -4/-4/-2 -3/-3/-2 -2/-2/-1 -1/-1/-1
o second
f 1/1/1 3/3/1 4/4/1
"""


def test_obj_first_mesh_left_handed(tmp_path):
    p = tmp_path / "m.obj"
    p.write_text(OBJ_TEXT)
    m = lib.load_obj(p)
    v = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0.5], [0, 1, -2]], np.float32)
    vt = np.array([[0, 0], [1, 0], [1, 0.25], [0, 1]], np.float32)
    vn = np.array([[0, 0, 1], [0, 1, 0]], np.float32)
    corners = [(0, 0, 0), (1, 1, 0), (2, 2, 1), (0, 0, 0), (1, 1, 0), (2, 2, 1), (3, 3, 1)]
    pos = np.array([v[c[0]] for c in corners]) * [1, 1, -1]
    uv = np.array([vt[c[1]] for c in corners])
    uv[:, 1] = 1 - uv[:, 1]
    nrm = np.array([vn[c[2]] for c in corners]) * [1, 1, -1]
    assert np.array_equal(m["positions"], pos.astype(np.float32))
    assert np.array_equal(m["uvs"], uv.astype(np.float32))
    assert np.array_equal(m["normals"], nrm.astype(np.float32))
    # triangle (0,1,2) reversed; quad (3,4,5,6) reversed to (6,5,4,3), fanned from 6
    assert m["indices"].tolist() == [[2, 1, 0], [6, 5, 4], [6, 4, 3]]


def test_obj_without_normals_or_uvs(tmp_path):
    p = tmp_path / "m.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nv 0 0 1\nvn 0 0 1\nf 1 2 3\nf 1//1 3//1 4//1\n")
    m = lib.load_obj(p)
    assert m["normals"] is None and m["uvs"] is None
    assert m["positions"].shape == (6, 3) and m["indices"].tolist() == [[2, 1, 0], [5, 4, 3]]
    assert m["positions"][5].tolist() == [0, 0, -1]


def test_obj_errors(tmp_path):
    for name, text in {"range": "v 0 0 0\nf 1 2 3\n", "empty": "v 0 0 0\n", "bad": "v 0 0 x\nf 1 1 1\n"}.items():
        p = tmp_path / (name + ".obj")
        p.write_text(text)
        with pytest.raises(lib.WpbrError):
            lib.load_obj(p)


# ------------------------------------------------------- oracle mesh bake ----

# world = M p + t with dyadic entries, so every product and sum is exact
M = np.array([[2, 1, 0], [0, 4, 0], [0, 0, 0.5]], np.float64)
T = np.array([1, -2, 3], np.float64)
MINV = np.array([[0.5, -0.125, 0], [0, 0.25, 0], [0, 0, 2]], np.float64)


def _xform():
    world = np.concatenate([M, T[:, None]], axis=1)
    inv = np.concatenate([MINV, (-MINV @ T)[:, None]], axis=1)
    assert np.array_equal(MINV @ M, np.eye(3))
    return np.concatenate([world.reshape(-1), inv.reshape(-1)]).astype(np.float32)


def test_oracle_bake_closed_form():
    rng = np.random.default_rng(1)
    pos = rng.integers(-64, 64, (50, 3)).astype(np.float32)
    nrm = rng.integers(-8, 8, (50, 3)).astype(np.float32)
    tan = rng.integers(-8, 8, (50, 3)).astype(np.float32)
    idx = rng.integers(0, 50, (80, 3)).astype(np.uint32)
    pw, nw, tw, tri = po.mesh_bake(_xform(), pos, idx, nrm, tan)
    assert np.array_equal(pw, (pos @ M.T + T).astype(np.float32))      # points: M p + t
    assert np.array_equal(tw, (tan @ M.T).astype(np.float32))          # vectors: no translation
    assert np.array_equal(nw, (nrm @ MINV).astype(np.float32))         # normals: Minv^T n
    assert not np.array_equal(nw, (nrm @ MINV.T).astype(np.float32))   # (not Minv n)
    assert np.array_equal(tri, pw[idx].reshape(-1, 9))


def test_oracle_bake_normals_stay_perpendicular():
    # a normal perpendicular to a surface tangent stays perpendicular under any affine map
    rng = np.random.default_rng(2)
    a = rng.normal(size=(3, 3))
    x = np.concatenate([np.concatenate([a, rng.normal(size=(3, 1))], 1).reshape(-1),
                        np.concatenate([np.linalg.inv(a), np.zeros((3, 1))], 1).reshape(-1)])
    tan = rng.normal(size=(200, 3)).astype(np.float32)
    nrm = np.cross(tan, rng.normal(size=(200, 3))).astype(np.float32)
    _, nw, tw, _ = po.mesh_bake(x, np.zeros((200, 3)), np.zeros((0, 3)), nrm, tan)
    cos = np.abs((nw * tw).sum(1)) / (np.linalg.norm(nw, axis=1) * np.linalg.norm(tw, axis=1))
    assert cos.max() < 1e-4


# ---------------------------------------------------- oracle texel decode ----


def test_oracle_srgb_decode_closed_form():
    b = np.arange(256, dtype=np.uint8)
    got = po.srgb8_to_linear(b)
    v = b.astype(np.float64) / 255.0
    ref = np.where(v <= 0.04045, v / 12.92, ((v + 0.055) / 1.055) ** 2.4)
    assert got[0] == 0.0 and got[255] == pytest.approx(1.0, abs=2e-7)
    assert np.all(np.diff(got) > 0)
    # linear segment up to byte 10 (10/255 = 0.0392 <= 0.04045 < 11/255)
    lin = (b[:11].astype(np.float32) / np.float32(255)) / np.float32(12.92)
    assert np.array_equal(got[:11], lin)
    # the sRGB decoding curve (IEC 61966-2-1) within float rounding of its terms
    assert np.max(np.abs(got - ref) / np.maximum(ref, 1e-30)) < 4e-7
    # textbook value: 50 % grey (byte 128) decodes to 0.2158605
    assert got[128] == pytest.approx(0.2158605, abs=2e-7)
