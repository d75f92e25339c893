"""Writes the committed fixtures of tests/golden/ from the oracle (run from the
repository root: python tests/golden/make_golden.py).  They freeze the oracle's
outputs on small seeded inputs so that a change of the oracle, or of the GPU
path, shows up against data that does not depend on building either.

  lbvh_small.npz   300 triangles + 2 spheres: the LBVH arrays (the reference's
                   key order) and the closest hits of 512 rays
  lbvh_small_full.npz  the same with the full 30-bit key order
  render_tiny.npz  the plain scene at 32x32, 1 spp: per-sample XYZ and the film
"""
import pathlib
import sys

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT))
import pyoracle as po  # noqa: E402
from woodenpbrengine_amd import scenes  # noqa: E402

HERE = pathlib.Path(__file__).resolve().parent


def lbvh_inputs():
    tris = scenes.random_triangles(300, 901, median_edge=0.05)
    c = np.array([[0.3, 0.4, 0.5], [0.7, 0.2, 0.6]])
    sx = np.stack([scenes.translate_xform(ci) for ci in c]).astype(np.float32)
    sr = np.array([0.05, 0.08], np.float32)
    rays = np.concatenate([scenes.pinhole_rays(16, 16, (0.5, 0.5, -1.0)), scenes.random_rays(256, 902)], axis=1)
    return sx, sr, tris, rays


def render_scene():
    return scenes.c0_scene(width=32, height=32, spp=1)


LBVH_FIXTURES = {po.ORDER_REFERENCE: "lbvh_small.npz", po.ORDER_FULL: "lbvh_small_full.npz"}


def main():
    sx, sr, tris, rays = lbvh_inputs()
    osc = po.OScene(sx, sr, tris)
    for order, name in LBVH_FIXTURES.items():
        ob = po.build_scene_bvh(osc, order=order)
        hit, t, _, ovf = po.collide(osc, ob, rays)
        assert not ovf
        out = {k: ob.get(k) for k in ("morton", "sorted_index", "node_offset", "node_nshapes", "node_bounds",
                                      "treelet_start")}
        ns = out["node_nshapes"]
        out["node_axis_interior"] = ob.get("node_axis")[ns == 0]
        np.savez_compressed(HERE / name, sphere_xform=sx, sphere_radius=sr, tri_verts=tris, rays=rays,
                            hit_shape=hit, hit_t=t, **out)
    r = po.render(render_scene(), nthreads=1)
    np.savez_compressed(HERE / "render_tiny.npz", sample_xyz=r["sample_xyz"], film_xyz=r["film_xyz"],
                        film_w=r["film_w"], film_rgb=r["film_rgb"])


if __name__ == "__main__":
    main()
