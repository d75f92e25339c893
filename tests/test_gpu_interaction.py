"""GPU parity of wpbr_scene_collide (runCollisionSystem with the interactions it
attaches: CSurfaceInteraction for bSurfInteraction rays, CInteraction for the
others, and the differentials of rays with a CRayDifferential) against the
oracle, through the C ABI.  Hit ids and t bit-exact; every interaction field
bit-exact."""
import numpy as np
import pytest

import pyoracle as po
from woodenpbrengine_amd import lib, scenes

pytestmark = pytest.mark.gpu


def _xform(A, c):
    A = np.asarray(A, np.float64)
    M = np.concatenate([A, np.asarray(c, np.float64)[:, None]], axis=1)
    Ai = np.linalg.inv(A)
    Mi = np.concatenate([Ai, (-Ai @ np.asarray(c, np.float64))[:, None]], axis=1)
    return np.concatenate([M.ravel(), Mi.ravel()]).astype(np.float32)


def _shapes_scene():
    """Spheres under rotations and anisotropic scales, a mesh with uvs whose
    determinant is 0 (the x0.001 reset, CSTriangle.cpp:204-219), the icosphere
    without attributes and the ground quad with uvs, normals and tangents."""
    g = np.random.default_rng(5)
    s = scenes.RenderScene()
    m = s.material(scenes.MAT_DIELECTRIC, alpha=0.1, eta=1.5)
    centres = [(-1.5, 0.2, 0.5), (1.4, -0.3, 0.2), (0.0, 1.3, 0.8)]
    for c in centres:
        s.add_sphere(c, 0.45, m)
    ico = scenes.icosphere(16)
    s.add_mesh(ico, m)
    tris = scenes.random_triangles(400, 6, median_edge=0.3) * np.float32(3) - np.float32(1.5)
    s.add_mesh(tris, m, np.full((len(tris), 6), 0.5, np.float32))
    qv, quv, qn, qt = scenes.ground_quad(12.0, 12.0, -1.0)
    s.add_mesh(qv, m, quv, qn, qt)
    s.add_light_sphere((0.0, 4.0, -1.0), 0.3, scenes.blackbody_spectrum(6500.0, 40.0))
    s.cam = scenes.camera(32, 32, (0.0, 0.9, -4.6), (0.0, -0.2, 0.0))
    s.width, s.height, s.spp, s.seed = 32, 32, 1, 1
    s = s.finalize()
    xf = []
    for c in centres:
        q, _ = np.linalg.qr(g.normal(size=(3, 3)))
        xf.append(_xform(q @ np.diag(g.uniform(0.5, 2.0, 3)), c))
    s.sphere_xform[:len(xf)] = np.stack(xf)
    return s


@pytest.mark.parametrize("which", ["c0", "shapes"])
def test_scene_collide_interactions_match_oracle(ctx, which):
    if which == "c0":
        s = scenes.c0_scene(width=64, height=64, spp=1)
        rays = np.concatenate([scenes.pinhole_rays(128, 64, (0.0, 0.9, -4.6)),
                               scenes.random_rays(8192, 31)], axis=1)
    else:
        s = _shapes_scene()
        rays = np.concatenate([scenes.pinhole_rays(128, 64, (0.0, 0.9, -6.0)),
                               scenes.random_rays(8192, 32)], axis=1)
        # random rays from inside the scene's box
        rays[:3, 8192:] = rays[:3, 8192:] * np.float32(4) - np.float32(2)
    rays = np.ascontiguousarray(rays, np.float32)
    n = rays.shape[1]
    g = np.random.default_rng(9)
    surf = (g.random(n) < 0.8).astype(np.uint8)
    # ray differentials: offset origins and slightly turned directions
    o, d = rays[:3], rays[3:]
    dx = d + g.normal(scale=1e-3, size=d.shape).astype(np.float32)
    dy = d + g.normal(scale=1e-3, size=d.shape).astype(np.float32)
    diff = np.concatenate([o + np.float32(1e-3), dx, o - np.float32(1e-3), dy]).astype(np.float32)
    print("scene %s: %d shapes, %d rays" % (which, s.n_shapes, n), flush=True)
    es, et, esi = po.scene_collide(s, rays, surf=surf, diff=diff)
    r = lib.Renderer(ctx, s)
    d_rays, d_surf, d_diff = ctx.array(rays), ctx.array(surf), ctx.array(diff)
    d_s, d_t, d_si = r.collide(d_rays, n, d_surf, d_diff)
    gs, gt, gsi = d_s.numpy(), d_t.numpy(), d_si.numpy()
    hit = gs != 0xFFFFFFFF
    print("hits %d (spheres %d), surface interactions %d" % (hit.sum(), (hit & (gs < len(s.sphere_radius))).sum(),
                                                              (hit & (surf == 1)).sum()), flush=True)
    np.testing.assert_array_equal(gs, es)
    np.testing.assert_array_equal(gt.view(np.uint32), et.view(np.uint32))
    for name, f in lib.SI_FIELDS.items():
        k = 2 if name == "uv" else (1 if name in ("time", "dudx", "dvdx", "dudy", "dvdy") else 3)
        a, b = gsi[f:f + k], esi[f:f + k]
        same = (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))
        assert same.all(), "%s: %d rays differ" % (name, (~same.all(axis=0)).sum())
    # the interaction without outputs: same hits
    d_s2, d_t2, _ = r.collide(d_rays, n, d_surf, None, with_si=False)
    np.testing.assert_array_equal(d_s2.numpy(), gs)
    r.destroy()
