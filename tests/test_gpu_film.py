"""GPU film output (wpbr_film_rgb8, JobOutputFilmTitles CSFilm.cpp:141-170) against
the oracle's tonemap: bit-exact bytes on the same film, for random films that
cover every branch (black, negative, linear and power segments, clamped, zero
weight sums) and for a rendered film; the PNG of a rendered film decodes to the
same bytes."""
import numpy as np
import pytest

import pyoracle as po
from woodenpbrengine_amd import lib, scenes
from test_film_output import _read_png

pytestmark = pytest.mark.gpu


def _random_film(rng, h, w):
    f = np.empty((h, w, 4), np.float32)
    # the weight sum sets the scale: v = rgb / (w 800000) spans ~1e-4 .. 2
    f[..., 3] = rng.uniform(0.2, 3.0, (h, w))
    f[..., :3] = (f[..., 3:4] * 800000.0) * rng.uniform(-0.3, 1.6, (h, w, 3)) \
        * np.exp(rng.uniform(-9.0, 0.0, (h, w, 1)))
    m = rng.random((h, w))
    f[m < 0.05, :3] = 0.0          # black pixels
    f[(m >= 0.05) & (m < 0.07), 3] = 0.0    # zero weight sums
    f[(m >= 0.07) & (m < 0.09), 3] *= -1.0  # negative weight sums (Mitchell lobes)
    return f


@pytest.mark.parametrize("h,w", [(1, 1), (37, 53), (512, 512)])
def test_film_rgb8_random(ctx, h, w):
    rng = np.random.default_rng(h * 7 + w)
    f = _random_film(rng, h, w)
    d = ctx.array(f)
    got = lib.film_rgb8(ctx, d, w, h)
    exp = po.film_rgb8(f[..., :3].reshape(-1, 3), f[..., 3].reshape(-1)).reshape(h, w, 3)
    d.free()
    print("film %dx%d: %d bytes differ, %d distinct values"
          % (h, w, int((got != exp).sum()), len(np.unique(got))), flush=True)
    assert np.array_equal(got, exp)


def test_film_rgb8_rendered(ctx, tmp_path):
    s = scenes.c0_scene(width=64, height=64, spp=2)
    r = lib.Renderer(ctx, s)
    r.render()
    f = r.d_film.numpy()
    got = r.film_rgb8()
    exp = po.film_rgb8(f[..., :3].reshape(-1, 3), f[..., 3].reshape(-1)).reshape(s.height, s.width, 3)
    assert np.array_equal(got, exp)
    # The reference's divisor (800000) suits its own light levels; this scene is
    # dimmer, so the film is also scaled into the full 8-bit range (the same
    # float32 scale on both sides) to exercise every branch on real radiance.
    y = f[..., 1] / np.where(f[..., 3] > 0, f[..., 3], 1.0)
    k = np.float32(0.8 * 800000.0 / np.percentile(y[f[..., 1] > 0], 99))
    fs = f.copy()
    fs[..., :3] *= k
    d = ctx.array(fs)
    got_s = lib.film_rgb8(ctx, d, s.width, s.height)
    d.free()
    exp_s = po.film_rgb8(fs[..., :3].reshape(-1, 3), fs[..., 3].reshape(-1)).reshape(got.shape)
    assert np.array_equal(got_s, exp_s)
    # against the oracle's own render: its film sums spectra, this one XYZ, so a
    # byte may step by one where a value sits on a truncation boundary
    ref = po.render(s)
    rs = ref["film_xyz"] * k
    eref = po.film_rgb8(rs.reshape(-1, 3), ref["film_w"].reshape(-1)).reshape(got.shape)
    diff = np.abs(got_s.astype(int) - eref.astype(int))
    print("rendered film (scale %.3g): %d distinct byte values, %d of %d bytes differ from the "
          "oracle render (max %d)" % (k, len(np.unique(got_s)), int((diff > 0).sum()), diff.size,
                                      int(diff.max())), flush=True)
    assert diff.max() <= 1 and (diff > 0).mean() < 1e-2
    p = tmp_path / "c0.png"
    r.write_png(p)
    assert np.array_equal(_read_png(p), got)
    r.destroy()
