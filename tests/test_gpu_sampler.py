"""Stratified sample arrays on the GPU (wpbr_sample_arrays,
JobSamplerStratifiedGenerateSampels1D/2D CSSamplerStratified.cpp:8-90) bit-exact
against the oracle's arrays, built the reference's way (stratified in loop
order, then shuffled); and the render with stratified samples against the
oracle's render of the same scene (per-sample XYZ within 1e-4 relative)."""
import numpy as np
import pytest

import pyoracle as po
from woodenpbrengine_amd import lib, scenes
from test_gpu_render import _compare

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("n1,n2", [(20, 32), (32, 80), (80, 272), (7, 100), (1, 1)])
def test_sample_arrays_match_oracle(ctx, mode, n1, n2):
    pixel0, npix = 1000, 4096
    g1, g2 = lib.sample_arrays(ctx, 20261203, mode, pixel0, npix, n1, n2)
    e1, e2 = po.sample_arrays(20261203, mode, pixel0, npix, n1, n2)
    print("mode %d n1 %d n2 %d: %d + %d values, %.3f ms" % (mode, n1, n2, g1.size, g2.size,
                                                          ctx.last_ms()), flush=True)
    assert np.array_equal(g1.view(np.uint32), e1.view(np.uint32))
    assert np.array_equal(g2.view(np.uint32), e2.view(np.uint32))


def test_sample_arrays_reject_bad_args(ctx):
    with pytest.raises(lib.WpbrError):
        lib.sample_arrays(ctx, 1, 3, 0, 4, 8, 8)
    with pytest.raises(lib.WpbrError):
        lib.sample_arrays(ctx, 1, 1, 0, 4, 0, 8)


@pytest.mark.parametrize("mode", [1, 2])
def test_render_stratified_matches_oracle(ctx, mode):
    s = scenes.c0_scene(width=64, height=64, spp=4)
    s.sampler = mode
    frac_bad, rel_rmse, exact = _compare(s, ctx)
    assert frac_bad == 0.0
    assert rel_rmse < 1e-5
