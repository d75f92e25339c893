"""Film output: the oracle's tonemap (JobOutputFilmTitles, CSFilm.cpp:141-170,
gammaCorrect CSFilm.cpp:129-140) against values fixed by hand, and the PNG
writer of the C ABI (host code, no GPU) decoded with zlib."""
import math
import struct
import zlib

import numpy as np
import pytest

import pyoracle as po
from woodenpbrengine_amd import lib

# xyzToRGB row of R (CoefficientSpectrum.h:16-21): X contributes 3.240479 to R.
R_OF_X = 3.240479


def _pixel_for_r(v):
    """XYZ = (1, 0, 0) and a weight sum giving R / (w * 800000) = v."""
    return np.array([[1.0, 0.0, 0.0]], np.float32), np.array([R_OF_X / (v * 800000.0)], np.float32)


def test_black_pixel_is_zero():
    out = po.film_rgb8(np.zeros((2, 3)), np.array([1.0, 0.0]))
    assert (out == 0).all()


def test_clamped_white_is_254():
    # 1.055f - 0.055f rounds to 0.99999994f (1 - 2^-24 is nearer than 1), and
    # 0.99999994f * 255 = 254.99998f, truncated to 254: a clamped channel never
    # reaches 255 in the reference's arithmetic.
    xyz, w = _pixel_for_r(50.0)
    out = po.film_rgb8(xyz, w)
    # G = -0.969256 X < 0 clamps to 0; B = 0.041556 X / d = 50 * 0.041556 / 3.240479 > 0.6
    assert out[0, 0] == 254 and out[0, 1] == 0
    assert abs(np.float32(1.055) - np.float32(0.055) - np.float32(1 - 2 ** -24)) == 0


def test_linear_segment_slope_12_98():
    # v = 0.00303: 12.98 v 255 = 10.029 -> 10 (sRGB's 12.92 would give 9.98 -> 9)
    xyz, w = _pixel_for_r(0.00303)
    out = po.film_rgb8(xyz, w)
    assert out[0, 0] == 10
    assert out[0, 1] == 0  # negative G clamps to 0


def test_power_segment_matches_srgb():
    # above 0.0031308 gammaCorrect is the sRGB transfer curve 1.055 v^(1/2.4) - 0.055
    rng = np.random.default_rng(7)
    vs = rng.uniform(0.004, 0.99, 400)
    for v in vs:
        xyz, w = _pixel_for_r(float(v))
        got = int(po.film_rgb8(xyz, w)[0, 0])
        vv = R_OF_X / float(w[0] * np.float32(800000.0))
        e = (1.055 * math.pow(vv, 1 / 2.4) - 0.055) * 255
        if abs(e - round(e)) < 1e-3:
            continue  # too near a step for a double-precision comparison
        assert got == int(e), (v, got, e)


def test_infinite_ratio_clamps():
    # a zero weight sum with non-zero radiance: R, B -> +inf -> 254, G -> -inf -> 0
    out = po.film_rgb8(np.array([[1.0, 0.0, 0.0]]), np.array([0.0]))
    assert out.tolist() == [[254, 0, 254]]


def _read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(data):
        (n,) = struct.unpack(">I", data[pos:pos + 4])
        typ, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(typ + body) & 0xFFFFFFFF
        chunks.append((typ, body))
        pos += 12 + n
    assert chunks[0][0] == b"IHDR" and chunks[-1][0] == b"IEND"
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (depth, ctype, comp, filt, inter) == (8, 2, 0, 0, 0)
    raw = zlib.decompress(b"".join(b for t, b in chunks if t == b"IDAT"))
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 3 * w)
    assert (rows[:, 0] == 0).all()
    return rows[:, 1:].reshape(h, w, 3)


@pytest.mark.parametrize("h,w", [(1, 1), (7, 5), (150, 200), (64, 64)])
def test_write_png_round_trip(tmp_path, h, w):
    img = np.random.default_rng(h * 1000 + w).integers(0, 256, (h, w, 3), dtype=np.uint8)
    p = tmp_path / "film.png"
    lib.write_png(p, img)
    assert np.array_equal(_read_png(p), img)


def test_write_png_rejects_bad_args(tmp_path):
    with pytest.raises(lib.WpbrError):
        lib.write_png(tmp_path / "x.png", np.zeros((0, 4, 3), np.uint8))
    with pytest.raises(lib.WpbrError):
        lib.write_png(tmp_path / "x.png", np.zeros((4, 4), np.uint8))
    with pytest.raises(lib.WpbrError):
        lib.write_png(tmp_path / "no_such_dir" / "x.png", np.zeros((4, 4, 3), np.uint8))
