"""Textured metals in the render on the GPU (CMaterialMetal.cpp:24-52 with the
camera's ray differentials, UV and sphere mapping and the EWA lookup),
through the C ABI: per-sample XYZ within 1e-4 relative of the oracle (0
samples out of bound) and the film within 1e-5 relative RMSE, for both
samplers; the job-by-job render and a batch with a textured scene equal the
plain render bit for bit; bad bindings and the differentiable render's
backward on a textured scene are errors."""
import numpy as np
import pytest

import pyoracle as po
from woodenpbrengine_amd import lib, scenes

pytestmark = pytest.mark.gpu

M_XYZ_RGB = np.array([[3.240479, -1.537150, -0.498535], [-0.969256, 1.875991, 0.041556],
                      [0.041556, -0.204043, 1.057311]])


@pytest.mark.parametrize("sampler", [0, 1])
def test_textured_render_matches_oracle(ctx, sampler):
    s = scenes.textured_scene(width=64, height=64, spp=2)
    s.sampler = sampler
    r = lib.Renderer(ctx, s)
    r.render(keep_samples=True)
    samp = r.d_samples.numpy().reshape(-1, 3)
    film = r.d_film.numpy().astype(np.float64)
    ref = po.render(s)
    exp = ref["sample_xyz"].reshape(-1, 3)
    bad = ~np.isclose(samp, exp, rtol=1e-4, atol=1e-6) & ~(np.isnan(samp) & np.isnan(exp))
    same = (samp.view(np.uint32) == exp.view(np.uint32)).all(axis=1)
    img = (film[..., :3] @ M_XYZ_RGB.T) / film[..., 3:4]
    eimg = ref["film_rgb"].astype(np.float64)
    ok = np.isfinite(img).all(-1) & np.isfinite(eimg).all(-1)
    rel = np.sqrt(np.mean((img[ok] - eimg[ok]) ** 2)) / np.sqrt(np.mean(eimg[ok] ** 2))
    print("samples %d, bit-exact %.4f, out of tolerance %d, film rmse / rms %.3e"
          % (len(samp), same.mean(), bad.any(axis=1).sum(), rel), flush=True)
    assert not bad.any()
    assert rel < 1e-5
    # the job sequence gives the same film
    f0 = r.d_film.numpy()
    r.render_staged()
    np.testing.assert_array_equal(r.d_film.numpy().view(np.uint32), f0.view(np.uint32))
    r.destroy()


def test_textured_batch_equals_single(ctx):
    ss = [scenes.textured_scene(width=64, height=64, spp=2), scenes.c0_scene(width=64, height=32, spp=2),
          scenes.textured_scene(width=32, height=64, spp=2, seed=77)]
    br = lib.BatchRenderer(ctx, ss)
    br.render()
    films = [br.film(i) for i in range(len(ss))]
    br.destroy()
    for i, s in enumerate(ss):
        r = lib.Renderer(ctx, s)
        r.render()
        np.testing.assert_array_equal(films[i].view(np.uint32), r.d_film.numpy().view(np.uint32), err_msg=str(i))
        r.destroy()


def test_texture_binding_errors(ctx):
    s = scenes.textured_scene(width=32, height=32, spp=1)
    bad = s.material_textures.copy()
    bad[0, 0] = len(s.textures)  # no such texture
    s.material_textures = bad
    with pytest.raises(lib.WpbrError):
        lib.Renderer(ctx, s)
    s = scenes.textured_scene(width=32, height=32, spp=1)
    diel = [m for m in range(len(s.material_table))
            if s.material_table[m, 0].view(np.uint32) == scenes.MAT_DIELECTRIC][0]
    s.material_textures[diel, 1] = 0  # a dielectric binding a texture
    with pytest.raises(lib.WpbrError):
        lib.Renderer(ctx, s)
    # the differentiable render's backward does not cover textured materials
    d = lib.DiffRenderer(ctx, scenes.textured_scene(width=32, height=32, spp=1))
    d.forward()
    d.set_target(np.zeros((32, 32, 3), np.float32))
    d.loss()
    with pytest.raises(lib.WpbrError):
        d.backward()
    d.destroy()
