"""Textures on the GPU (wpbr_texture_create / _level / _lookup) bit-exact
against the oracle, square and rectangular (one-texel-wide levels included):
every MIP level (built on the device), and anisotropic (EWA) lookups over random
points and footprints from magnification to the coarsest level, anisotropy
beyond the 32:1 clamp, zero footprints (bilinear path) and points on the
texture's edges."""
import numpy as np
import pytest

import pyoracle as po
from woodenpbrengine_amd import lib

pytestmark = pytest.mark.gpu


def _queries(rng, n):
    q = np.empty((n, 6), np.float32)
    q[:, :2] = rng.random((n, 2))
    scale = np.exp(rng.uniform(-9.0, 0.5, (n, 1)))
    q[:, 2:4] = rng.normal(0, 1, (n, 2)) * scale
    aniso = np.exp(rng.uniform(-5.0, 1.0, (n, 1)))  # minor/major from 1/150 to 2.7
    q[:, 4:6] = rng.normal(0, 1, (n, 2)) * scale * aniso
    q[: n // 20, 4:6] = 0.0              # zero minor axis: bilinear at level 0
    q[n // 20: n // 10, :2] = rng.integers(0, 2, (n // 10 - n // 20, 2))  # corners
    return q


@pytest.mark.parametrize("wh", [(4, 4), (8, 8), (64, 64), (512, 512), (64, 16), (17, 100), (256, 4), (100, 1)])
def test_texture_matches_oracle(ctx, wh):
    w, h = wh
    rng = np.random.default_rng(w * 1000 + h)
    rgb = rng.random((h, w, 3), dtype=np.float32) ** 2.2
    t = lib.Texture(ctx, rgb)
    exp_levels = po.texture_levels(rgb)
    assert t.n_levels == len(exp_levels)
    for l, e in enumerate(exp_levels):
        assert np.array_equal(t.level(l).view(np.uint32), e.view(np.uint32)), l
    q = _queries(rng, 20000)
    got = t.lookup(q)
    exp = po.texture_lookup(rgb, q)
    bad = ~(got.view(np.uint32) == exp.view(np.uint32)).all(axis=1)
    print("%dx%d: %d levels, %d lookups, %d differ, %.3f ms" % (w, h, t.n_levels, len(q), int(bad.sum()),
                                                                ctx.last_ms()), flush=True)
    t.destroy()
    assert not bad.any(), q[bad][:5]


def test_texture_rejects_bad_sizes(ctx):
    # the longer side below 4 (no level), or a side beyond 16384
    for h, w in ((2, 2), (3, 1), (1, 20000)):
        with pytest.raises(lib.WpbrError):
            lib.Texture(ctx, np.zeros((h, w, 3), np.float32))
