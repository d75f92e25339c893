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
