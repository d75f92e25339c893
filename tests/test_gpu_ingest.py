"""Scene ingestion on the GPU through the C ABI, bit-exact against the oracle:
the world-space mesh bake (wpbr_mesh_bake, STriangleMesh::create,
CSTriangleMesh.cpp:11-39) on small and 2M-triangle meshes (de-indexed as assimp
gives them, and indexed with random gathers), the index check, the texel decode
(wpbr_srgb8_to_linear, CSTextureProcess.h:322-329) of every byte value and of an
env-map-sized image, and OBJ file -> bake -> LBVH -> closest hits, PNG file ->
texel decode -> texture."""
import numpy as np
import pytest

import pyoracle as po
from woodenpbrengine_amd import lib, scenes

pytestmark = pytest.mark.gpu


def _affine(rng):
    """A random world 3x4 with its inverse (24 floats)."""
    a = rng.normal(size=(3, 3)) + 2 * np.eye(3)
    t = rng.normal(size=3)
    ai = np.linalg.inv(a)
    return np.concatenate([np.concatenate([a, t[:, None]], 1).reshape(-1),
                           np.concatenate([ai, (-ai @ t)[:, None]], 1).reshape(-1)]).astype(np.float32)


def _bake_both(ctx, x, pos, idx, nrm=None, tan=None):
    d_pos, d_idx = ctx.array(pos), ctx.array(idx)
    d_nrm = ctx.array(nrm) if nrm is not None else None
    d_tan = ctx.array(tan) if tan is not None else None
    d_tri, d_pw, d_nw, d_tw = lib.mesh_bake(ctx, x, d_pos, len(pos), d_idx, len(idx), d_nrm, d_tan)
    got = [d_tri.numpy(), d_pw.numpy(), d_nw.numpy() if d_nw else None, d_tw.numpy() if d_tw else None]
    for a in (d_pos, d_idx, d_nrm, d_tan, d_tri, d_pw, d_nw, d_tw):
        if a is not None:
            a.free()
    pw, nw, tw, tri = po.mesh_bake(x, pos, idx, nrm, tan)
    return got, [tri, pw, nw, tw]


def _same(a, b):
    return a is None and b is None or np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_bake_small_all_attributes(ctx):
    rng = np.random.default_rng(11)
    pos = rng.normal(size=(3000, 3)).astype(np.float32)
    nrm = rng.normal(size=(3000, 3)).astype(np.float32)
    tan = rng.normal(size=(3000, 3)).astype(np.float32)
    idx = rng.integers(0, 3000, (5001, 3)).astype(np.uint32)
    got, want = _bake_both(ctx, _affine(rng), pos, idx, nrm, tan)
    for g, w in zip(got, want):
        assert _same(g, w)


@pytest.mark.parametrize("indexed", [False, True])
def test_bake_two_million_triangles(ctx, indexed):
    rng = np.random.default_rng(12)
    nt = 2_009_780
    if indexed:   # shared vertices, random gathers
        pos = rng.uniform(-1, 1, (nt // 2, 3)).astype(np.float32)
        idx = rng.integers(0, len(pos), (nt, 3)).astype(np.uint32)
    else:         # one vertex per corner, as assimp hands an OBJ mesh over
        pos = scenes.random_triangles(nt, 5).reshape(-1, 3)
        idx = np.arange(3 * nt, dtype=np.uint32).reshape(-1, 3)
    print("bake", "indexed" if indexed else "de-indexed", nt, flush=True)
    got, want = _bake_both(ctx, _affine(rng), pos, idx)
    assert _same(got[0], want[0]) and _same(got[1], want[1])
    if not indexed:  # the gather of a de-indexed mesh is the baked vertex array itself
        assert np.array_equal(got[0].reshape(-1, 3), got[1])


def test_bake_rejects_index_out_of_range(ctx):
    x = _affine(np.random.default_rng(1))
    pos = np.zeros((4, 3), np.float32)
    d_pos, d_bad = ctx.array(pos), ctx.array(np.array([[0, 1, 4]], np.uint32))
    with pytest.raises(lib.WpbrError):
        lib.mesh_bake(ctx, x, d_pos, 4, d_bad, 1)
    d_pos.free()
    d_bad.free()
    # the next call starts from a clear error count
    got, want = _bake_both(ctx, x, pos, np.array([[0, 1, 3]], np.uint32))
    assert _same(got[0], want[0])


def test_srgb_decode_every_byte_and_env_map(ctx):
    rng = np.random.default_rng(13)
    for data in (np.arange(256, dtype=np.uint8), rng.integers(0, 256, 3 * 1024 * 512, dtype=np.uint8)):
        d_in = ctx.array(data)
        d_out = lib.srgb8_to_linear(ctx, d_in, data.size)
        assert _same(d_out.numpy(), po.srgb8_to_linear(data))
        d_in.free()
        d_out.free()


def test_obj_to_bake_to_closest_hits(ctx, tmp_path):
    tris = scenes.icosphere(8)  # 1280 triangles, written de-indexed
    path = tmp_path / "sphere.obj"
    lines = ["v %.9g %.9g %.9g" % tuple(v) for v in tris.reshape(-1, 3)]
    lines += ["f %d %d %d" % (3 * i + 1, 3 * i + 2, 3 * i + 3) for i in range(len(tris))]
    path.write_text("\n".join(lines) + "\n")
    m = lib.load_obj(path)
    assert m["positions"].shape == (3 * len(tris), 3)
    x = _affine(np.random.default_rng(14))
    d_pos, d_idx = ctx.array(m["positions"]), ctx.array(m["indices"])
    d_tri, d_pw, _, _ = lib.mesh_bake(ctx, x, d_pos, len(m["positions"]), d_idx, len(m["indices"]))
    tri = d_tri.numpy()
    for a in (d_pos, d_idx, d_tri, d_pw):
        a.free()
    _, _, _, otri = po.mesh_bake(x, m["positions"], m["indices"])
    assert _same(tri, otri)
    sx = np.stack([scenes.translate_xform((4.0, 0.0, 0.0))])
    sr = np.array([0.5], np.float32)
    rays = np.concatenate([scenes.pinhole_rays(64, 32, (0.0, 0.0, -8.0)), scenes.random_rays(4096, 9)],
                          axis=1).astype(np.float32)
    sc = lib.Scene(ctx, sx, sr, tri)
    bvh = lib.Lbvh(ctx)
    bvh.build(sc)
    d_rays = ctx.array(rays)
    d_s, d_t, _ = lib.collide(ctx, sc, bvh, d_rays, rays.shape[1])
    osc = po.OScene(sx, sr, otri)
    os_, ot, _, _ = po.collide(osc, po.build_scene_bvh(osc), rays)
    assert np.array_equal(d_s.numpy(), os_)
    assert _same(d_t.numpy(), ot)
    assert (os_ != 0xFFFFFFFF).sum() > 500
    bvh.destroy()


def test_png_to_texture(ctx, tmp_path):
    rng = np.random.default_rng(15)
    img = rng.integers(0, 256, (64, 64, 3), dtype=np.uint8)
    lib.write_png(tmp_path / "tex.png", img)
    rgb8 = lib.read_png(tmp_path / "tex.png")
    d_in = ctx.array(rgb8)
    d_lin = lib.srgb8_to_linear(ctx, d_in, rgb8.size)
    lin = d_lin.numpy().reshape(64, 64, 3)
    d_in.free()
    d_lin.free()
    tex = lib.Texture(ctx, lin)
    assert _same(tex.level(0), po.srgb8_to_linear(img))
    tex.destroy()
