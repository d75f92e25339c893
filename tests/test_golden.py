"""Committed fixtures (tests/golden/, written by make_golden.py from the oracle):
the oracle must still reproduce them exactly, and the GPU path must match them
(LBVH arrays and hits bit for bit, per-sample radiance within 1e-4 relative)."""
import pathlib

import numpy as np
import pytest

import pyoracle as po
from woodenpbrengine_amd import lib, scenes

G = pathlib.Path(__file__).resolve().parent / "golden"


ORDERS = {po.ORDER_REFERENCE: "lbvh_small.npz", po.ORDER_FULL: "lbvh_small_full.npz"}


def _lbvh(order):
    return np.load(G / ORDERS[order])


@pytest.mark.parametrize("order", list(ORDERS))
def test_oracle_reproduces_lbvh_fixture(order):
    f = _lbvh(order)
    osc = po.OScene(f["sphere_xform"], f["sphere_radius"], f["tri_verts"])
    ob = po.build_scene_bvh(osc, order=order)
    for k in ("morton", "sorted_index", "node_offset", "node_nshapes", "node_bounds", "treelet_start"):
        np.testing.assert_array_equal(ob.get(k), f[k], err_msg=k)
    hit, t, _, _ = po.collide(osc, ob, f["rays"])
    np.testing.assert_array_equal(hit, f["hit_shape"])
    np.testing.assert_array_equal(t.view(np.uint32), f["hit_t"].view(np.uint32))


def test_oracle_reproduces_render_fixture():
    f = np.load(G / "render_tiny.npz")
    r = po.render(scenes.c0_scene(width=32, height=32, spp=1), nthreads=2)
    np.testing.assert_array_equal(r["sample_xyz"].view(np.uint32), f["sample_xyz"].view(np.uint32))
    np.testing.assert_array_equal(r["film_w"], f["film_w"])


@pytest.mark.gpu
@pytest.mark.parametrize("order", list(ORDERS))
def test_gpu_matches_lbvh_fixture(ctx, order):
    f = _lbvh(order)
    sc = lib.Scene(ctx, f["sphere_xform"], f["sphere_radius"], f["tri_verts"])
    b = lib.Lbvh(ctx, order)
    b.build(sc)
    for k in ("morton", "sorted_index", "node_offset", "node_nshapes", "node_bounds", "treelet_start"):
        np.testing.assert_array_equal(b.download(k), f[k], err_msg=k)
    ns = b.download("node_nshapes")
    np.testing.assert_array_equal(b.download("node_axis")[ns == 0], f["node_axis_interior"])
    d_rays = ctx.array(f["rays"])
    d_s, d_t, _ = lib.collide(ctx, sc, b, d_rays, f["rays"].shape[1])
    np.testing.assert_array_equal(d_s.numpy(), f["hit_shape"])
    np.testing.assert_array_equal(d_t.numpy().view(np.uint32), f["hit_t"].view(np.uint32))
    b.destroy()


@pytest.mark.gpu
def test_gpu_matches_render_fixture(ctx):
    f = np.load(G / "render_tiny.npz")
    r = lib.Renderer(ctx, scenes.c0_scene(width=32, height=32, spp=1))
    r.render(keep_samples=True)
    got = r.d_samples.numpy()
    assert np.isclose(got, f["sample_xyz"], rtol=1e-4, atol=1e-6).all()
    r.destroy()
