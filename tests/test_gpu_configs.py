"""The build+traversal microbenchmark, the full render and the batched
many-scene render at their full sizes on the GPU (through the C ABI).

* 1 048 576 triangles: the LBVH arrays in both key orders bit-exact against the
  oracle; the closest hits of every 64th of the 4 194 304 rays (65 536 rays,
  pinhole and random) bit-exact against the oracle; all 4 194 304 rays give the
  same t on both trees (a closest hit does not depend on the tree) and the same
  primitive except on exact t ties.
* 1024x1024 x 16 spp (16 777 216 samples, 2 009 850 shapes, the reference's
  stratified sampler): per-sample XYZ within 1e-4 relative and the film within
  1e-5 relative RMSE on 16 film tiles (every 64th) against the oracle; over the
  whole film every pixel has a positive filter weight, at most 1e-5 of the
  samples are non-finite, and the film is the Mitchell-weighted sum of the
  samples it holds (checked through the image mean of the oracle's tiles).
* 64 scenes, 8 spp: every film of one batched render equals that scene's
  single-scene render bit for bit.
"""
import numpy as np
import pytest

import pyoracle as po
from woodenpbrengine_amd import lib, scenes

pytestmark = pytest.mark.gpu

M_XYZ_RGB = np.array([[3.240479, -1.537150, -0.498535], [-0.969256, 1.875991, 0.041556],
                      [0.041556, -0.204043, 1.057311]])


def _film_rgb(f):
    f = f.astype(np.float64)
    return (f[..., :3] @ M_XYZ_RGB.T) / f[..., 3:4]


def test_c1_build_and_traversal_full_size(ctx):
    tris, rays = scenes.microbench()
    n = rays.shape[1]
    print("microbench: %d triangles, %d rays" % (len(tris), n), flush=True)
    sx, sr = np.zeros((0, 24), np.float32), np.zeros(0, np.float32)
    osc = po.OScene(sx, sr, tris)
    sc = lib.Scene(ctx, sx, sr, tris)
    d_rays = ctx.array(rays)
    sub = np.arange(0, n, 64)
    hits = {}
    for order in (lib.ORDER_REFERENCE, lib.ORDER_FULL):
        ob = po.build_scene_bvh(osc, order=order)
        b = lib.Lbvh(ctx, order)
        b.build(sc)
        c = b.counts()
        print("order %d: %d nodes, %d treelets" % (order, c["nodes"], c["treelets"]), flush=True)
        assert c["nodes"] == ob.n_nodes and c["treelets"] == ob.n_treelets
        for f in ("morton", "sorted_code", "sorted_index", "treelet_start", "node_nshapes", "node_offset",
                  "node_bounds"):
            np.testing.assert_array_equal(b.download(f), ob.get(f), err_msg=f)
        ns = b.download("node_nshapes")
        np.testing.assert_array_equal(b.download("node_axis")[ns == 0], ob.get("node_axis")[ns == 0])
        d_s, d_t, _ = lib.collide(ctx, sc, b, d_rays, n)
        gs, gt = d_s.numpy(), d_t.numpy()
        os_, ot, _, ovf = po.collide(osc, ob, np.ascontiguousarray(rays[:, sub]))
        assert not ovf
        np.testing.assert_array_equal(gs[sub], os_)
        np.testing.assert_array_equal(gt[sub].view(np.uint32), ot.view(np.uint32))
        hits[order] = (gs, gt)
        b.destroy()
    (s0, t0), (s1, t1) = hits[lib.ORDER_REFERENCE], hits[lib.ORDER_FULL]
    np.testing.assert_array_equal(t0.view(np.uint32), t1.view(np.uint32))
    differ = s0 != s1
    print("rays whose primitive differs between the trees (t ties): %d" % differ.sum(), flush=True)
    assert differ.sum() <= 1e-4 * n
    miss = s0 == 0xFFFFFFFF
    print("misses: %d primary, %d random" % (miss[:n // 2].sum(), miss[n // 2:].sum()), flush=True)
    # random rays start inside the cube: with 1M triangles of ~2e-3 edge the mean
    # free path is ~1.2 (one over triangles per volume times their mean
    # projected area), about twice the mean distance to the cube's boundary
    assert 0.3 < miss[n // 2:].mean() < 0.8


def test_c2_render_full_size(ctx):
    s = scenes.c2_scene()
    s.sampler = lib.SAMPLER_STRATIFIED
    W, H, spp = s.width, s.height, s.spp
    print("full render: %dx%d x %d spp, %d shapes" % (W, H, spp, s.n_shapes), flush=True)
    r = lib.Renderer(ctx, s)
    r.render(keep_samples=True)
    samp = r.d_samples.numpy().reshape(H, W, spp, 3)
    film = r.d_film.numpy().reshape(H, W, 4)
    r.destroy()
    # whole-film properties
    assert (film[..., 3] > 0).all() and np.isfinite(film[..., 3]).all()
    bad = ~np.isfinite(samp).all(axis=-1)
    print("non-finite samples: %d of %d" % (bad.sum(), bad.size), flush=True)
    assert bad.mean() <= 1e-5
    img = _film_rgb(film)
    fin = np.isfinite(img).all(axis=-1)
    assert fin.mean() > 0.999 and (img[fin].mean(axis=0) > 0).all()
    # 16 tiles (every 64th) against the oracle
    tiles = np.arange(0, (W // 32) * (H // 32), 64, dtype=np.uint32)
    ref = po.render(s, tiles=tiles)
    exp = ref["sample_xyz"].reshape(H, W, spp, 3)
    eimg = ref["film_rgb"].astype(np.float64)
    worst, n_bad, n_tot, se, ss = 0.0, 0, 0, 0.0, 0.0
    for t in tiles:
        ty, tx = divmod(int(t), W // 32)
        ys, xs = slice(32 * ty, 32 * ty + 32), slice(32 * tx, 32 * tx + 32)
        g, e = samp[ys, xs], exp[ys, xs]
        ok = ~np.isclose(g, e, rtol=1e-4, atol=1e-6)
        both_nan = np.isnan(g) & np.isnan(e)
        n_bad += (ok & ~both_nan).any(axis=-1).sum()
        n_tot += g.shape[0] * g.shape[1] * g.shape[2]
        gi, ei = img[ys, xs], eimg[ys, xs]
        m = np.isfinite(gi).all(axis=-1) & np.isfinite(ei).all(axis=-1)
        se += float(((gi[m] - ei[m]) ** 2).sum())
        ss += float((ei[m] ** 2).sum())
    rel = np.sqrt(se / ss)
    print("tiles %d, samples %d, out of tolerance %d, film rmse / rms %.3e" % (len(tiles), n_tot, n_bad, rel),
          flush=True)
    assert n_bad == 0
    assert rel < 1e-5


def test_c3_batch_full_size(ctx):
    ss = scenes.c3_scenes()
    print("batch: %d scenes, %d shapes, %d samples" % (len(ss), sum(s.n_shapes for s in ss),
                                                        sum(s.width * s.height * s.spp for s in ss)), flush=True)
    br = lib.BatchRenderer(ctx, ss)
    br.render()
    films = [br.film(i) for i in range(len(ss))]
    st = br.stats()
    assert st["samples"] == sum(s.width * s.height * s.spp for s in ss)
    br.destroy()
    for i, s in enumerate(ss):
        r = lib.Renderer(ctx, s)
        r.render()
        single = r.d_film.numpy()
        r.destroy()
        np.testing.assert_array_equal(films[i].view(np.uint32), single.view(np.uint32),
                                      err_msg="scene %d (%dx%d, %d shapes)" % (i, s.width, s.height, s.n_shapes))
