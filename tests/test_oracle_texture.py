"""The oracle's texture filtering (CSTextureProcess.h:134-345,
CSTextureProcess.cpp:16-32) against values fixed by hand: the Gaussian table's
ends, MIP levels of constant and linear textures (including the reference's
last-column / last-row rule), exact texel values through the bilinear path,
and EWA lookups of linear ramps at points where every level's texel grid is
symmetric (a filter centred wrongly, or scaled by the wrong level size, moves
the result)."""
import math

import numpy as np
import pytest

import pyoracle as po


def ramp(res, a=1.0, b=0.0, c=0.0):
    """texel (x, y) = a x + b y + c, as res x res x 3 (channels scaled 1, 2, 3)."""
    y, x = np.mgrid[0:res, 0:res].astype(np.float64)
    v = a * x + b * y + c
    return np.stack([v, 2 * v, 3 * v], -1).astype(np.float32)


def test_filter_table():
    w = po.texture_filter()
    assert abs(w[0] - (1 - math.exp(-2))) < 1e-7
    assert w[63] == 0.0
    assert (np.diff(w) < 0).all()
    assert abs(w[21] - (math.exp(-2 * 21 / 63) - math.exp(-2))) < 1e-7


@pytest.mark.parametrize("res", [4, 8, 64, 256])
def test_mip_levels(res):
    lv = po.texture_levels(np.full((res, res, 3), 0.75, np.float32))
    assert len(lv) == int(math.log2(res)) - 1
    assert all((l == 0.75).all() for l in lv)
    # texel (x, y) = x + 64 y: a 2x2 mean of level l-1 is the value at its centre
    lv = po.texture_levels(ramp(res, 1.0, 64.0))
    for l in range(1, len(lv)):
        r, prev = res >> l, lv[l - 1][..., 0].astype(np.float64)
        cur = lv[l][..., 0]
        y, x = np.mgrid[0:r - 1, 0:r - 1]
        np.testing.assert_array_equal(cur[:r - 1, :r - 1],
                                      (prev[2 * y, 2 * x] + prev[2 * y, 2 * x + 1] + prev[2 * y + 1, 2 * x]
                                       + prev[2 * y + 1, 2 * x + 1]) / 4)
        # last column: the previous level's last column only; corner copied
        yy = np.arange(r - 1)
        np.testing.assert_array_equal(cur[yy, r - 1], (prev[2 * yy, -1] + prev[2 * yy + 1, -1]) / 2)
        np.testing.assert_array_equal(cur[r - 1, yy], (prev[-1, 2 * yy] + prev[-1, 2 * yy + 1]) / 2)
        assert cur[r - 1, r - 1] == prev[-1, -1]


@pytest.mark.parametrize("wh", [(64, 16), (17, 100), (256, 4), (5, 3), (100, 1), (3, 40)])
def test_mip_levels_rectangular(wh):
    """generateMIPS on W x H (CSTextureProcess.h:349-371): (uint16_t)(log2(max
    side) - 1) levels, each max(1, half) per axis; inside, 2x2 means of a
    linear ramp; the last column / row from the previous level's last column /
    row; the last texel copied -- also where a level is one texel wide."""
    w, h = wh
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    rgb = np.stack([x + 64 * y, 2 * (x + 64 * y), 3 * (x + 64 * y)], -1).astype(np.float32)
    lv = po.texture_levels(rgb)
    assert len(lv) == int(math.log2(max(w, h))) - 1
    pw, ph = w, h
    np.testing.assert_array_equal(lv[0], rgb)
    for l in range(1, len(lv)):
        rw, rh = max(1, pw // 2), max(1, ph // 2)
        assert lv[l].shape == (rh, rw, 3)
        prev, cur = lv[l - 1][..., 0].astype(np.float64), lv[l][..., 0]
        yy, xx = np.mgrid[0:rh - 1, 0:rw - 1]
        np.testing.assert_array_equal(cur[:rh - 1, :rw - 1],
                                      (prev[2 * yy, 2 * xx] + prev[2 * yy, 2 * xx + 1] + prev[2 * yy + 1, 2 * xx]
                                       + prev[2 * yy + 1, 2 * xx + 1]) / 4)
        ys, xs = np.arange(rh - 1), np.arange(rw - 1)
        np.testing.assert_array_equal(cur[ys, rw - 1], (prev[2 * ys, pw - 1] + prev[2 * ys + 1, pw - 1]) / 2)
        np.testing.assert_array_equal(cur[rh - 1, xs], (prev[ph - 1, 2 * xs] + prev[ph - 1, 2 * xs + 1]) / 2)
        assert cur[rh - 1, rw - 1] == prev[ph - 1, pw - 1]
        pw, ph = rw, rh


def test_rectangular_lookup_scales_per_axis():
    """On a W x H texture the EWA scales s by W and t by H: a texel centre
    ((x + 0.5) / W, (y + 0.5) / H) with a zero footprint reads texel (x, y)
    exactly, and a small isotropic footprint on a ramp along s alone averages
    to the ramp's value there."""
    w, h = 64, 16
    rng = np.random.default_rng(4)
    tex = rng.random((h, w, 3), dtype=np.float32)
    xs, ys = rng.integers(0, w, 40), rng.integers(0, h, 40)
    q = np.zeros((40, 6), np.float32)
    q[:, 0] = (xs + 0.5) / w
    q[:, 1] = (ys + 0.5) / h
    np.testing.assert_array_equal(po.texture_lookup(tex, q), tex[ys, xs])
    ramp_s = np.repeat(np.arange(w, dtype=np.float32)[None, :, None], h, 0).repeat(3, 2)
    q = np.array([[0.5, 0.5, 1e-3, 0, 0, 1e-3]], np.float32)
    np.testing.assert_allclose(po.texture_lookup(ramp_s, q)[0], 0.5 * w - 0.5, rtol=1e-6)


def test_zero_footprint_reads_texels():
    res = 16
    rng = np.random.default_rng(3)
    tex = rng.random((res, res, 3), dtype=np.float32)
    xs, ys = rng.integers(0, res, 50), rng.integers(0, res, 50)
    q = np.zeros((50, 6), np.float32)
    q[:, 0], q[:, 1] = (xs + 0.5) / res, (ys + 0.5) / res
    q[:, 2] = rng.random(50)  # dstdx alone: minor = |dstdy| = 0 -> bilinear at level 0
    out = po.texture_lookup(tex, q)
    np.testing.assert_array_equal(out, tex[ys, xs])


# footprints that stay clear of the coarse levels' last row and column, whose
# texels come from the previous level's edge only and break a ramp
@pytest.mark.parametrize("res,width", [(16, 0.5), (16, 3.0), (64, 0.5), (64, 3.0), (64, 12.0),
                                       (256, 0.5), (256, 3.0), (256, 12.0)])
def test_ewa_of_a_ramp_at_symmetric_points(res, width):
    # every level's texel grid is symmetric about (0.5, 0.5); an isotropic or
    # axis-aligned anisotropic footprint there averages a ramp to its value.
    # The third footprint's dstdy is the longer axis: the reference never swaps
    # the axes, so its length picks the level.
    tex = ramp(res, 1.0 / res, 0.37 / res, 0.1)
    ratio = 8 if width <= 3 else 2
    for dx, dy in [((width, 0), (0, width)), ((width * ratio, 0), (0, width)), ((0, width), (width * 2, 0))]:
        q = np.array([[0.5, 0.5, dx[0] / res, dx[1] / res, dy[0] / res, dy[1] / res]], np.float32)
        got = po.texture_lookup(tex, q)[0]
        exp = (0.5 - 0.5 / res) * (1 + 0.37) + 0.1
        np.testing.assert_allclose(got, [exp, 2 * exp, 3 * exp], rtol=2e-5)


def test_constant_texture_is_constant_everywhere():
    res = 64
    tex = np.full((res, res, 3), 0.3, np.float32)
    rng = np.random.default_rng(9)
    q = np.empty((300, 6), np.float32)
    q[:, :2] = rng.random((300, 2))
    q[:, 2:] = rng.normal(0, 1, (300, 4)) * np.exp(rng.uniform(-8, 0, (300, 1)))
    out = po.texture_lookup(tex, q)
    np.testing.assert_allclose(out, 0.3, rtol=1e-5)


def test_anisotropy_clamp_picks_the_level_only():
    """Beyond 32:1 the clamp lengthens the minor axis for the level choice only
    (CSTextureProcess.h:151-157); ewa() reads the caller's dstdx / dstdy again
    (:183-185).  A clamped lookup equals the lerp of the two levels' EWA with the
    unscaled derivatives, at the levels the clamped length picks, and differs
    from EWA with the scaled minor axis."""
    res = 64
    rng = np.random.default_rng(21)
    tex = rng.random((res, res, 3), dtype=np.float32)
    n_levels = len(po.texture_levels(tex))
    q = []
    for k in range(40):
        ang = rng.uniform(0, np.pi)
        # lod = n_levels - 1 + log2(major / 32) at k / 10 + 0.05 (levels 0 .. n_levels - 2)
        lod_target = (k % (10 * (n_levels - 1))) / 10 + 0.05
        major = 32 * 2.0 ** (lod_target - (n_levels - 1))
        minor = major / rng.uniform(60.0, 400.0)
        u = np.array([np.cos(ang), np.sin(ang)])
        v = np.array([-np.sin(ang), np.cos(ang)])
        q.append([rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.8), *(major * u), *(minor * v)])
    q = np.array(q, np.float32)
    got = po.texture_lookup(tex, q)
    worst_scaled = 0.0
    for i in range(len(q)):
        major = np.hypot(np.float64(q[i, 2]), np.float64(q[i, 3]))
        lod = max(0.0, n_levels - 1 + np.log2(major / 32))
        il = int(np.floor(lod))
        assert il < n_levels - 1 and 1e-2 < lod - il < 1 - 1e-2
        d = lod - il
        a = po.texture_ewa(tex, il, q[i:i + 1])[0].astype(np.float64)
        b = po.texture_ewa(tex, il + 1, q[i:i + 1])[0].astype(np.float64)
        np.testing.assert_allclose(got[i], (1 - d) * a + d * b, rtol=2e-6, atol=1e-7)
        # the ellipse with the minor derivative scaled up to major / 32 is another filter
        qs = q[i:i + 1].copy()
        qs[0, 4:] *= np.float32(major / 32 / np.hypot(np.float64(q[i, 4]), np.float64(q[i, 5])))
        sa = po.texture_ewa(tex, il, qs)[0].astype(np.float64)
        sb = po.texture_ewa(tex, il + 1, qs)[0].astype(np.float64)
        worst_scaled = max(worst_scaled, float(np.abs((1 - d) * sa + d * sb - got[i]).max()))
    assert worst_scaled > 1e-3
