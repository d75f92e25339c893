"""GPU parity: LBVH build and closest-hit traversal against the oracle, in both
key orders (the reference's bits-0..23 order, the default, and the full sort).

Bit-exact: Morton codes, sort permutation, every node array (bounds compared as
float32 values), hit shape ids; t equal bit for bit (the north-star bound is
2 ULP).  Calls go through the C ABI (libwpbr_hip.so)."""
import numpy as np
import pytest

import pyoracle as po
from woodenpbrengine_amd import lib, scenes

pytestmark = pytest.mark.gpu


def _spheres(k, seed):
    g = np.random.default_rng(seed)
    c = g.random((k, 3))
    r = (0.02 + 0.08 * g.random(k)).astype(np.float32)
    sx = np.stack([scenes.translate_xform(ci) for ci in c]) if k else np.zeros((0, 24), np.float32)
    return sx, r


def _case(name):
    if name == "tiny":
        tris = scenes.random_triangles(5, 11)
        sx, sr = _spheres(1, 12)
    elif name == "small":
        tris = scenes.random_triangles(3000, 13, median_edge=0.02)
        sx, sr = _spheres(3, 14)
    elif name == "medium":
        tris = scenes.random_triangles(100_000, 15, median_edge=0.01)
        sx, sr = _spheres(8, 16)
    elif name == "icosphere":
        tris = scenes.icosphere(32)
        sx, sr = _spheres(2, 17)
    elif name == "clustered":
        # many coincident centroids: duplicate Morton codes and >= 30-shape leaves
        base = scenes.random_triangles(50, 18, median_edge=0.05)
        tris = np.repeat(base, 80, axis=0)
        sx, sr = _spheres(0, 19)
    elif name == "runs":
        # groups of 1..60 coincident triangles among unique ones: leaves of 33..64
        # equal codes, and runs of equal codes inside leaves whose codes differ
        g = np.random.default_rng(31)
        base = scenes.random_triangles(300, 32, median_edge=0.05)
        tris = np.concatenate([np.repeat(base, g.integers(1, 61, len(base)), axis=0),
                               scenes.random_triangles(2000, 33, median_edge=0.02)])
        sx, sr = _spheres(2, 34)
    else:
        raise KeyError(name)
    return sx, sr, tris


CASES = ["tiny", "small", "medium", "icosphere", "clustered", "runs"]
# (block, LDS stack depth, mapping, node width, refill, chunk); stack depth 8
# forces spills; mapping 3 is the persistent dynamic-fetch kernel
TRACE_CONFIGS = [(128, 32, 1, 2, 16, 256), (64, 16, 1, 4, 16, 256), (256, 8, 2, 4, 16, 256),
                 (64, 8, 1, 2, 16, 256), (64, 16, 3, 4, 16, 256), (128, 8, 3, 2, 1, 64),
                 (64, 16, 3, 4, 64, 4096)]
DEFAULT_TRACE = (64, 16, 1, 4, 16, 256)


ORDERS = [lib.ORDER_REFERENCE, lib.ORDER_FULL]


@pytest.mark.parametrize("order", ORDERS)
@pytest.mark.parametrize("name", CASES)
def test_lbvh_build_bit_exact(ctx, name, order):
    sx, sr, tris = _case(name)
    print("case", name, len(sr), len(tris), "order", order, flush=True)
    osc = po.OScene(sx, sr, tris)
    ob = po.build_scene_bvh(osc, order=order)
    sc = lib.Scene(ctx, sx, sr, tris)
    b = lib.Lbvh(ctx, order)
    b.build(sc)
    c = b.counts()
    assert c["nodes"] == ob.n_nodes
    assert c["treelets"] == ob.n_treelets
    np.testing.assert_array_equal(b.download("centroid_bounds").ravel(),
                                  ob.get("centroid_bounds").ravel())
    np.testing.assert_array_equal(b.download("morton"), ob.get("morton"))
    np.testing.assert_array_equal(b.download("sorted_code"), ob.get("sorted_code"))
    np.testing.assert_array_equal(b.download("sorted_index"), ob.get("sorted_index"))
    np.testing.assert_array_equal(b.download("treelet_start"), ob.get("treelet_start"))
    ns = b.download("node_nshapes")
    np.testing.assert_array_equal(ns, ob.get("node_nshapes"))
    np.testing.assert_array_equal(b.download("node_offset"), ob.get("node_offset"))
    interior = ns == 0
    np.testing.assert_array_equal(b.download("node_axis")[interior], ob.get("node_axis")[interior])
    np.testing.assert_array_equal(b.download("node_bounds"), ob.get("node_bounds"))
    b.destroy()


@pytest.mark.parametrize("order", ORDERS)
@pytest.mark.parametrize("name", CASES)
def test_collide_matches_oracle(ctx, name, order):
    sx, sr, tris = _case(name)
    print("case", name, "order", order, flush=True)
    osc = po.OScene(sx, sr, tris)
    ob = po.build_scene_bvh(osc, order=order)
    rays = np.concatenate([scenes.pinhole_rays(128, 64, (0.5, 0.5, -1.0)),
                           scenes.random_rays(8192, 21)], axis=1)
    os_, ot, _, ovf = po.collide(osc, ob, rays)
    assert not ovf
    sc = lib.Scene(ctx, sx, sr, tris)
    b = lib.Lbvh(ctx, order)
    b.build(sc)
    d_rays = ctx.array(rays)
    try:
        # every traversal configuration gives the same hits
        for cfg in TRACE_CONFIGS:
            ctx.set_trace_config(*cfg)
            d_s, d_t, _ = lib.collide(ctx, sc, b, d_rays, rays.shape[1])
            gs, gt = d_s.numpy(), d_t.numpy()
            np.testing.assert_array_equal(gs, os_, err_msg="config %s" % (cfg,))
            np.testing.assert_array_equal(gt.view(np.uint32), ot.view(np.uint32),
                                          err_msg="config %s" % (cfg,))
    finally:
        ctx.set_trace_config(*DEFAULT_TRACE)
    b.destroy()


def test_rebuild_reuses_buffers(ctx):
    """Builds in a row (smaller, then larger, with the key order switched
    between them) give the oracle's trees."""
    b = lib.Lbvh(ctx)
    assert b.order == lib.ORDER_REFERENCE
    for n, seed, order in ((20000, 31, lib.ORDER_REFERENCE), (700, 32, lib.ORDER_FULL),
                           (700, 32, lib.ORDER_REFERENCE), (30000, 33, lib.ORDER_FULL)):
        lib.load().wpbr_lbvh_set_order(b.h, order)
        tris = scenes.random_triangles(n, seed, median_edge=0.02)
        sx, sr = _spheres(2, seed)
        sc = lib.Scene(ctx, sx, sr, tris)
        b.build(sc)
        ob = po.build_scene_bvh(po.OScene(sx, sr, tris), order=order)
        assert b.counts()["treelets"] == ob.n_treelets
        np.testing.assert_array_equal(b.download("node_offset"), ob.get("node_offset"))
        np.testing.assert_array_equal(b.download("node_bounds"), ob.get("node_bounds"))
    b.destroy()


def test_autotune_collide(ctx):
    """The autotuner keeps one of its candidates and the hits stay the oracle's."""
    sx, sr, tris = _case("medium")
    osc = po.OScene(sx, sr, tris)
    ob = po.build_scene_bvh(osc)
    rays = scenes.random_rays(20000, 23)
    os_, ot, _, ovf = po.collide(osc, ob, rays)
    sc = lib.Scene(ctx, sx, sr, tris)
    b = lib.Lbvh(ctx)
    b.build(sc)
    d_rays = ctx.array(rays)
    try:
        cfg = lib.autotune_collide(ctx, sc, b, d_rays, rays.shape[1])
        assert cfg == ctx.trace_config()
        assert cfg["mapping"] in (1, 3) and cfg["width"] in (2, 4) and cfg["block"] in (64, 128)
        d_s, d_t, _ = lib.collide(ctx, sc, b, d_rays, rays.shape[1])
        np.testing.assert_array_equal(d_s.numpy(), os_)
        np.testing.assert_array_equal(d_t.numpy().view(np.uint32), ot.view(np.uint32))
    finally:
        ctx.set_trace_config(*DEFAULT_TRACE)
    b.destroy()


@pytest.mark.parametrize("order", ORDERS)
def test_staged_build_matches_and_checks_order(ctx, order):
    """wpbr_lbvh_build_stage: the six jobs in order equal wpbr_lbvh_build; a
    stage out of order is rejected."""
    sx, sr, tris = _case("small")
    sc = lib.Scene(ctx, sx, sr, tris)
    b = lib.Lbvh(ctx, order)
    with pytest.raises(lib.WpbrError):
        b.build_stage(sc, 3)
    for stage in range(6):
        b.build_stage(sc, stage)
    ob = po.build_scene_bvh(po.OScene(sx, sr, tris), order=order)
    np.testing.assert_array_equal(b.download("node_offset"), ob.get("node_offset"))
    np.testing.assert_array_equal(b.download("node_bounds"), ob.get("node_bounds"))
    b.build_stage(sc, 0)
    with pytest.raises(lib.WpbrError):
        b.build_stage(sc, 2)
    b.destroy()


def _oracle_compare(b, ob):
    c = b.counts()
    assert c["nodes"] == ob.n_nodes and c["treelets"] == ob.n_treelets
    for f in ("morton", "sorted_code", "sorted_index", "treelet_start", "node_nshapes", "node_offset",
              "node_bounds"):
        np.testing.assert_array_equal(b.download(f), ob.get(f), err_msg=f)
    ns = b.download("node_nshapes")
    np.testing.assert_array_equal(b.download("node_axis")[ns == 0], ob.get("node_axis")[ns == 0])


@pytest.mark.parametrize("order", ORDERS)
def test_segmented_build_matches_oracle(ctx, order):
    """wpbr_lbvh_build_segmented: every scene's tree equals the oracle's (and so
    the single-scene build's), for scenes from 1 to 100 000 shapes in one batch;
    a second batch of other sizes reuses the same objects; collision on the
    segmented trees gives the oracle's hits; scenes of two key orders in one
    batch are refused."""
    cases = [_case(n) for n in CASES]
    cases.append((np.zeros((0, 24), np.float32), np.zeros(0, np.float32), scenes.random_triangles(1, 41)))
    sx1, sr1 = _spheres(1, 42)
    cases.append((sx1, sr1, np.zeros((0, 9), np.float32)))
    for k, n in enumerate((40, 900, 5000)):
        cases.append((*_spheres(k, 43 + k), scenes.random_triangles(n, 50 + k, median_edge=0.03)))
    scs = [lib.Scene(ctx, *c) for c in cases]
    bvhs = [lib.Lbvh(ctx, order) for _ in cases]
    print("segmented build of %d scenes" % len(scs), flush=True)
    lib.build_segmented(ctx, scs, bvhs)
    rays = np.concatenate([scenes.pinhole_rays(32, 32, (0.5, 0.5, -1.0)), scenes.random_rays(2048, 44)], axis=1)
    d_rays = ctx.array(rays)
    for i, (c, sc, b) in enumerate(zip(cases, scs, bvhs)):
        print("scene", i, sc.n_shapes, flush=True)
        osc = po.OScene(*c)
        ob = po.build_scene_bvh(osc, order=order)
        _oracle_compare(b, ob)
        os_, ot, _, ovf = po.collide(osc, ob, rays)
        d_s, d_t, _ = lib.collide(ctx, sc, b, d_rays, rays.shape[1])
        np.testing.assert_array_equal(d_s.numpy(), os_)
        np.testing.assert_array_equal(d_t.numpy().view(np.uint32), ot.view(np.uint32))
    # second batch: other scenes (some larger, some smaller) into the same objects
    cases2 = [(*_spheres(2, 60 + k), scenes.random_triangles(n, 70 + k, median_edge=0.02))
              for k, n in enumerate((30000, 3, 700))]
    scs2 = [lib.Scene(ctx, *c) for c in cases2]
    lib.build_segmented(ctx, scs2, bvhs[:3])
    for c, b in zip(cases2, bvhs[:3]):
        _oracle_compare(b, po.build_scene_bvh(po.OScene(*c), order=order))
    with pytest.raises(lib.WpbrError):
        lib.build_segmented(ctx, scs2[:2], [bvhs[0], bvhs[0]])
    other = lib.Lbvh(ctx, 1 - order)
    with pytest.raises(lib.WpbrError):
        lib.build_segmented(ctx, scs2[:2], [bvhs[0], other])
    other.destroy()
    for b in bvhs:
        b.destroy()
