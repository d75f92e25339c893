"""Batched many-scene render (wpbr_scenes_create + wpbr_batch_*) through the C
ABI: every scene's film equals the single-scene render bit for bit, graph
replays and a recaptured graph (another traversal configuration) give the same
films, and one scene's image matches the oracle's as the single-scene render
parity does (film RMSE below 1e-5 of the image RMS)."""
import numpy as np
import pytest

import pyoracle as po
from woodenpbrengine_amd import lib, scenes

pytestmark = pytest.mark.gpu


def _small_batch(spp=2):
    # scenes of the batched workload's kind, fewer triangles and samples
    ss = scenes.c3_scenes(n_scenes=6, spp=spp, max_triangles=20_000)
    # plus the plain scene at another size
    ss.append(scenes.c0_scene(width=64, height=96, spp=spp))
    if spp % 2:
        # scenes with stratified sample arrays in the same batch
        ss[1].sampler = 1
        ss[-1].sampler = 2
    return ss


# passes alternate between two streams: an odd count ends on the other one
@pytest.mark.parametrize("spp", [2, 3])
def test_batch_films_equal_single_renders(ctx, spp):
    ss = _small_batch(spp)
    print("batch of", [(s.width, s.height, s.n_shapes) for s in ss], flush=True)
    br = lib.BatchRenderer(ctx, ss)
    br.render()
    films = [br.film(i) for i in range(len(ss))]
    st = br.stats()
    assert st["samples"] == sum(s.width * s.height * s.spp for s in ss)
    assert st["secondary_rays"] > 0
    for i, s in enumerate(ss):
        r = lib.Renderer(ctx, s)
        r.render()
        single = r.d_film.numpy()
        r.destroy()
        print("scene", i, s.width, s.height, s.n_shapes, flush=True)
        np.testing.assert_array_equal(films[i].view(np.uint32), single.view(np.uint32))
    # replays of the captured graph
    br.render()
    br.render()
    for i in range(len(ss)):
        np.testing.assert_array_equal(br.film(i).view(np.uint32), films[i].view(np.uint32))
    # another traversal configuration: the graph is captured again
    saved = ctx.trace_config()
    try:
        ctx.set_trace_config(block=128, stack_depth=8, width=2)
        br.render()
        for i in range(len(ss)):
            np.testing.assert_array_equal(br.film(i).view(np.uint32), films[i].view(np.uint32))
    finally:
        ctx.set_trace_config(**saved)
    br.destroy()


def test_batch_scene_matches_oracle(ctx):
    s0 = scenes.c0_scene(width=64, height=64, spp=2)
    ss = [scenes.c3_scenes(n_scenes=1, spp=2, max_triangles=5_000)[0], s0]
    br = lib.BatchRenderer(ctx, ss)
    br.render()
    f = br.film(1).astype(np.float64)
    M = np.array([[3.240479, -1.537150, -0.498535], [-0.969256, 1.875991, 0.041556],
                  [0.041556, -0.204043, 1.057311]])
    img = (f[..., :3] @ M.T) / f[..., 3:4]
    ref = po.render(s0)["film_rgb"].astype(np.float64)
    rel = np.sqrt(np.mean((img - ref) ** 2)) / np.sqrt(np.mean(ref ** 2))
    print("film rmse / rms %.3e" % rel, flush=True)
    assert rel < 1e-5
    br.destroy()


def test_batch_autotune_keeps_films(ctx):
    ss = _small_batch()[:3]
    br = lib.BatchRenderer(ctx, ss)
    br.render()
    ref = [br.film(i) for i in range(len(ss))]
    saved = ctx.trace_config()
    try:
        cfg = br.autotune()
        assert cfg == ctx.trace_config() and cfg["mapping"] == 1
        br.render()
        for i in range(len(ss)):
            np.testing.assert_array_equal(br.film(i).view(np.uint32), ref[i].view(np.uint32))
    finally:
        ctx.set_trace_config(**saved)
    br.destroy()


def test_batch_c3_scenes_match_oracle(ctx):
    """Scenes of the batched workload (displaced icospheres from per-scene seeds,
    64 spheres, 4 sphere lights and an env map, the full render's material mix,
    films of the batch's sizes) rendered in one batch: each film equals the
    oracle's render of that scene (film RMSE below 1e-5 of the image RMS)."""
    ss = scenes.c3_scenes(n_scenes=4, spp=2, max_triangles=30_000)
    br = lib.BatchRenderer(ctx, ss)
    br.render()
    M = np.array([[3.240479, -1.537150, -0.498535], [-0.969256, 1.875991, 0.041556],
                  [0.041556, -0.204043, 1.057311]])
    for i, s in enumerate(ss):
        f = br.film(i).astype(np.float64)
        img = (f[..., :3] @ M.T) / f[..., 3:4]
        ref = po.render(s)["film_rgb"].astype(np.float64)
        ok = np.isfinite(ref).all(axis=-1) & np.isfinite(img).all(axis=-1)
        rel = np.sqrt(np.mean((img[ok] - ref[ok]) ** 2)) / np.sqrt(np.mean(ref[ok] ** 2))
        print("scene %d: %dx%d, %d shapes, film rmse / rms %.3e, non-finite pixels %d / %d"
              % (i, s.width, s.height, s.n_shapes, rel, (~ok).sum(), (~np.isfinite(ref).all(axis=-1)).sum()),
              flush=True)
        assert rel < 1e-5
        # the same pixels are non-finite on both sides
        np.testing.assert_array_equal(np.isfinite(img).all(axis=-1), np.isfinite(ref).all(axis=-1))
    br.destroy()
