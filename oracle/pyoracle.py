"""ctypes wrapper of oracle/_build/liboracle.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, as the checker and the timed CPU baseline; never by
the product package.
"""
import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"

_lib = None


def build():
    subprocess.run(["make", "-C", str(HERE), "-s"], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.exists():
        build()
    lib = C.CDLL(str(LIB))
    vp, u32, f32 = C.c_void_p, C.c_uint32, C.c_float
    lib.orc_shape_bounds.argtypes = [u32, vp, vp, u32, vp, vp, vp]
    lib.orc_centroid_bounds.argtypes = [u32, vp, vp]
    lib.orc_lbvh_build.argtypes = [u32, vp, vp, vp, C.c_int]
    lib.orc_lbvh_build.restype = vp
    lib.orc_lbvh_free.argtypes = [vp]
    lib.orc_lbvh_info.argtypes = [vp, vp]
    lib.orc_lbvh_copy.argtypes = [vp, C.c_int, vp]
    lib.orc_lbvh_copy.restype = C.c_int
    lib.orc_collide.argtypes = [u32, vp, vp, u32, vp, vp, u32, vp, vp, vp, C.c_int, vp]
    lib.orc_collide.restype = C.c_uint64
    lib.orc_brute_force.argtypes = [u32, vp, vp, u32, vp, u32, vp, vp, vp, C.c_int]
    lib.orc_tri_intersect.argtypes = [vp, vp, vp, vp]
    lib.orc_tri_intersect.restype = f32
    lib.orc_sphere_intersect.argtypes = [vp, f32, vp, vp]
    lib.orc_sphere_intersect.restype = f32
    lib.orc_box_hit.argtypes = [vp, vp, vp, f32, vp]
    lib.orc_box_hit.restype = C.c_int
    lib.orc_partition_check.argtypes = [u32, vp, C.c_int, vp, vp, vp]
    _lib = lib
    return lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a.reshape(shape) if shape is not None else a


def default_threads():
    return int(os.environ.get("WPBR_ORACLE_THREADS", os.cpu_count() or 1))


class OScene:
    def __init__(self, sphere_xform, sphere_radius, tri_verts):
        self.sx = _f32(sphere_xform, (-1, 24))
        self.sr = _f32(sphere_radius, (-1,))
        self.tv = _f32(tri_verts, (-1, 9))
        self.ns, self.nt = len(self.sr), len(self.tv)

    def args(self):
        return (self.ns, _p(self.sx), _p(self.sr), self.nt, _p(self.tv))


def shape_bounds(sc):
    lib = load()
    n = sc.ns + sc.nt
    b = np.empty((n, 6), np.float32)
    c = np.empty((n, 3), np.float32)
    lib.orc_shape_bounds(*sc.args(), _p(b), _p(c))
    return b, c


# Key orders of the LBVH build (oracle.h): the reference's (what its sort job
# leaves in shapeMortonCodes: bits 0..23, stable) or the complete 30-bit sort.
ORDER_REFERENCE = 0
ORDER_FULL = 1


class OLbvh:
    def __init__(self, bounds, centroid, entities=None, order=ORDER_REFERENCE):
        lib = load()
        self.lib = lib
        self.n = len(bounds)
        self.bounds = _f32(bounds, (-1, 6))
        self.centroid = _f32(centroid, (-1, 3))
        ent = None if entities is None else np.ascontiguousarray(entities, np.uint32)
        self.order = order
        self.h = lib.orc_lbvh_build(self.n, _p(self.bounds), _p(self.centroid), _p(ent), order)
        info = np.zeros(3, np.uint32)
        lib.orc_lbvh_info(self.h, _p(info))
        self.n_nodes, self.n_sah_nodes, self.n_treelets = (int(x) for x in info)

    def get(self, name):
        spec = {"morton": (0, self.n, np.uint32), "sorted_code": (1, self.n, np.uint32),
                "sorted_index": (2, self.n, np.uint32), "shape_entities": (3, self.n, np.uint32),
                "node_offset": (4, self.n_nodes, np.uint32),
                "node_nshapes": (5, self.n_nodes, np.uint32),
                "node_axis": (6, self.n_nodes, np.uint8),
                "node_bounds": (7, 6 * self.n_nodes, np.float32),
                "treelet_start": (8, self.n_treelets, np.uint32),
                "centroid_bounds": (9, 6, np.float32)}[name]
        out = np.empty(spec[1], dtype=spec[2])
        self.lib.orc_lbvh_copy(self.h, spec[0], _p(out))
        if name in ("node_bounds", "centroid_bounds"):
            out = out.reshape(-1, 6)
        return out

    def __del__(self):
        try:
            self.lib.orc_lbvh_free(self.h)
        except Exception:
            pass


def build_scene_bvh(sc, order=ORDER_REFERENCE):
    b, c = shape_bounds(sc)
    return OLbvh(b, c, order=order)


def collide(sc, bvh, rays, nthreads=None):
    lib = load()
    rays = _f32(rays, (6, -1))
    n = rays.shape[1]
    shape = np.empty(n, np.uint32)
    t = np.empty(n, np.float32)
    ovf = np.zeros(1, np.int32)
    cand = lib.orc_collide(*sc.args(), bvh.h, n, _p(rays), _p(shape), _p(t),
                           nthreads or default_threads(), _p(ovf))
    return shape, t, int(cand), bool(ovf[0])


def brute_force(sc, rays, nthreads=None):
    lib = load()
    rays = _f32(rays, (6, -1))
    n = rays.shape[1]
    shape = np.empty(n, np.uint32)
    t = np.empty(n, np.float32)
    lib.orc_brute_force(*sc.args(), n, _p(rays), _p(shape), _p(t), nthreads or default_threads())
    return shape, t


def tri_intersect(v9, o, d):
    lib = load()
    v9, o, d = _f32(v9), _f32(o), _f32(d)
    bary = np.zeros(3, np.float32)
    t = lib.orc_tri_intersect(_p(v9), _p(o), _p(d), _p(bary))
    return float(t), bary


def sphere_intersect(xform24, radius, o, d):
    lib = load()
    return float(lib.orc_sphere_intersect(_p(_f32(xform24)), float(radius), _p(_f32(o)),
                                          _p(_f32(d))))


def box_hit(b6, o, d, tmax=np.inf):
    lib = load()
    tn = np.zeros(1, np.float32)
    h = lib.orc_box_hit(_p(_f32(b6)), _p(_f32(o)), _p(_f32(d)), float(tmax), _p(tn))
    return bool(h), float(tn[0])


def partition_check(buckets, best):
    lib = load()
    b = np.ascontiguousarray(buckets, np.uint8)
    n = len(b)
    a1 = np.empty(n, np.uint32)
    a2 = np.empty(n, np.uint32)
    mids = np.empty(2, np.uint32)
    lib.orc_partition_check(n, _p(b), int(best), _p(a1), _p(a2), _p(mids))
    return a1, a2, mids


# ---- render -------------------------------------------------------------------
class SceneDesc(C.Structure):
    _fields_ = [("n_spheres", C.c_uint32), ("sphere_xform", C.c_void_p),
                ("sphere_radius", C.c_void_p), ("n_triangles", C.c_uint32),
                ("tri_verts", C.c_void_p), ("tri_flags", C.c_void_p), ("tri_uv", C.c_void_p),
                ("tri_normal", C.c_void_p), ("tri_tangent", C.c_void_p),
                ("shape_material", C.c_void_p), ("n_materials", C.c_uint32),
                ("materials", C.c_void_p), ("n_sphere_lights", C.c_uint32),
                ("light_shape", C.c_void_p), ("light_emit", C.c_void_p),
                ("has_env", C.c_uint32), ("env_emit", C.c_void_p), ("env_rgb", C.c_void_p),
                ("env_w", C.c_uint32), ("env_h", C.c_uint32), ("env_xform", C.c_void_p),
                ("raster_camera", C.c_void_p), ("camera_world", C.c_void_p),
                ("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32),
                ("seed", C.c_uint32), ("sampler", C.c_uint32),
                ("n_textures", C.c_uint32), ("tex_size", C.c_void_p), ("tex_rgb", C.c_void_p),
                ("material_textures", C.c_void_p)]


def scene_desc(s):
    """Host-pointer scene description of a woodenpbrengine_amd.scenes.RenderScene.
    Returns (desc, keepalive)."""
    keep = [np.ascontiguousarray(a) for a in (
        s.sphere_xform, s.sphere_radius, s.tri_verts, s.tri_flags, s.tri_uv, s.tri_normal,
        s.tri_tangent, s.shape_material, s.material_table, s.light_shape, s.light_emit,
        s.env_emit, s.env_rgb, s.env_xform, s.raster_camera, s.camera_world)]
    (sx, sr, tv, tf, tuv, tn, tt, sm, mt, ls, le, ee, er, ex, rc, cw) = keep
    texs = [np.ascontiguousarray(t, np.float32) for t in getattr(s, "textures", [])]
    sizes = np.array([[t.shape[1], t.shape[0]] for t in texs], np.uint32).reshape(-1, 2)
    ptrs = (C.c_void_p * max(1, len(texs)))(*[_p(t) for t in texs])
    mtex = np.ascontiguousarray(getattr(s, "material_textures", np.full((len(mt), 2), -1)), np.int32)
    keep += texs + [sizes, ptrs, mtex]
    d = SceneDesc(len(sr), _p(sx), _p(sr), len(tv), _p(tv), _p(tf), _p(tuv), _p(tn), _p(tt),
                  _p(sm), len(mt), _p(mt), len(ls), _p(ls), _p(le),
                  1 if s.env is not None else 0, _p(ee), _p(er), er.shape[1], er.shape[0],
                  _p(ex), _p(rc), _p(cw), s.width, s.height, s.spp, s.seed,
                  getattr(s, "sampler", 0), len(texs), _p(sizes), C.cast(ptrs, C.c_void_p), _p(mtex))
    return d, keep


SI_FLOATS = 49


def scene_collide(s, rays, surf=None, diff=None, nthreads=None):
    """runCollisionSystem on a render scene's geometry: closest hit shape and t
    per ray, and the interaction each ray cast gets (SI_FLOATS x n SoA floats,
    the layout of wpbr.h's WPBR_SI_*)."""
    lib = load()
    lib.orc_scene_collide.argtypes = [C.c_void_p, C.c_int, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p]
    d, keep = scene_desc(s)
    rays = _f32(rays, (6, -1))
    n = rays.shape[1]
    sf = None if surf is None else np.ascontiguousarray(surf, np.uint8)
    df = None if diff is None else _f32(diff, (12, n))
    shape = np.empty(n, np.uint32)
    t = np.empty(n, np.float32)
    si = np.empty((SI_FLOATS, n), np.float32)
    lib.orc_scene_collide(C.byref(d), nthreads or default_threads(), n, _p(rays), _p(sf), _p(df), _p(shape),
                          _p(t), _p(si))
    del keep
    return shape, t, si


def bsdf_eval(s, sample, frame, wo, material, wi=None, u=None):
    """The BSDF Compute (sample False) or Sample (True) jobs for requests of the
    scene's materials (SoA: frame 9 x n, wo 3 x n, wi 3 x n, u 2 x n): value
    n x 64 (f |cos wi|, a request's spectrum contiguous), pdf n, and for Sample
    wi 3 x n (world)."""
    lib = load()
    vp = C.c_void_p
    lib.orc_bsdf_eval.argtypes = [vp, C.c_int, C.c_uint32, vp, vp, vp, vp, vp, vp, vp, vp]
    d, keep = scene_desc(s)
    frame = _f32(frame, (9, -1))
    n = frame.shape[1]
    wo = _f32(wo, (3, n))
    wi = None if wi is None else _f32(wi, (3, n))
    u = None if u is None else _f32(u, (2, n))
    mat = np.ascontiguousarray(material, np.uint32)
    value = np.empty((n, 64), np.float32)
    pdf = np.empty(n, np.float32)
    wi_out = np.empty((3, n), np.float32)
    lib.orc_bsdf_eval(C.byref(d), 1 if sample else 0, n, _p(frame), _p(wo), _p(wi), _p(u), _p(mat), _p(value),
                      _p(pdf), _p(wi_out))
    del keep
    return (value, pdf, wi_out) if sample else (value, pdf)


def render(s, nthreads=None, tiles=None):
    """Render a scene on the CPU; returns a dict of numpy arrays."""
    lib = load()
    lib.orc_render.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_uint32]
    lib.orc_render.restype = C.c_void_p
    lib.orc_render_copy.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    lib.orc_render_free.argtypes = [C.c_void_p]
    d, keep = scene_desc(s)
    t = None if tiles is None else np.ascontiguousarray(tiles, np.uint32)
    h = lib.orc_render(C.byref(d), nthreads or default_threads(), _p(t), 0 if t is None else len(t))
    W, H, spp = s.width, s.height, s.spp
    out = {"sample_xyz": np.empty((H * W * spp, 3), np.float32),
           "sample_pfilm": np.empty((H * W * spp, 2), np.float32),
           "film_rgb": np.empty((H, W, 3), np.float32),
           "film_xyz": np.empty((H, W, 3), np.float32),
           "film_w": np.empty((H, W), np.float32),
           "counters": np.empty(8, np.uint64)}
    for i, k in enumerate(["sample_xyz", "sample_pfilm", "film_rgb", "film_xyz", "film_w",
                           "counters"]):
        lib.orc_render_copy(h, i, _p(out[k]))
    lib.orc_render_free(h)
    del keep
    return out


def spectrum_tables():
    lib = load()
    lib.orc_spectrum_tables.argtypes = [C.c_void_p]
    out = np.empty(192 + 2 * 448, np.float32)
    lib.orc_spectrum_tables(_p(out))
    return {"X": out[:64], "Y": out[64:128], "Z": out[128:192],
            "refl": out[192:640].reshape(7, 64), "illum": out[640:].reshape(7, 64)}


def from_rgb(rgb, illuminant=False):
    lib = load()
    lib.orc_from_rgb.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    out = np.empty(64, np.float32)
    lib.orc_from_rgb(_p(_f32(rgb)), 1 if illuminant else 0, _p(out))
    return out


def to_xyz(spec):
    lib = load()
    lib.orc_to_xyz.argtypes = [C.c_void_p, C.c_void_p]
    out = np.empty(3, np.float32)
    lib.orc_to_xyz(_p(_f32(spec)), _p(out))
    return out


def film_rgb8(xyz, w):
    """8-bit RGB of a film: xyz (npix, 3) accumulated XYZ, w (npix,) filter weight sums."""
    lib = load()
    lib.orc_film_rgb8.argtypes = [C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]
    xyz = _f32(xyz).reshape(-1, 3)
    w = _f32(w).reshape(-1)
    assert len(w) == len(xyz)
    out = np.empty((len(w), 3), np.uint8)
    lib.orc_film_rgb8(len(w), _p(xyz), _p(w), _p(out))
    return out


def sample_arrays(seed, mode, pixel0, npix, n1, n2, by_element=False):
    """Per-pixel sample arrays (npix x n1, npix x n2 x 2) of a sampler mode."""
    lib = load()
    lib.orc_sample_arrays.argtypes = [C.c_uint32] * 6 + [C.c_int, C.c_void_p, C.c_void_p]
    u1 = np.empty((npix, n1), np.float32)
    u2 = np.empty((npix, n2, 2), np.float32)
    lib.orc_sample_arrays(seed, mode, pixel0, npix, n1, n2, 1 if by_element else 0, _p(u1), _p(u2))
    return u1, u2


def permute_index(c, n, key):
    lib = load()
    lib.orc_permute_index.argtypes = [C.c_uint32] * 3
    lib.orc_permute_index.restype = C.c_uint32
    return int(lib.orc_permute_index(c, n, key))


def texture_levels(rgb):
    """MIP levels of an RGB texture (height x width x 3): list of h_l x w_l x 3 arrays."""
    lib = load()
    lib.orc_texture_levels.argtypes = [C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.orc_texture_levels.restype = C.c_uint32
    rgb = _f32(rgb)
    h, w = rgb.shape[:2]
    sizes = np.zeros(32, np.uint32)
    n = int(lib.orc_texture_levels(w, h, _p(rgb), _p(sizes), None))
    out = np.empty(sum(int(sizes[2 * l]) * int(sizes[2 * l + 1]) * 3 for l in range(n)), np.float32)
    lib.orc_texture_levels(w, h, _p(rgb), None, _p(out))
    levels, o = [], 0
    for l in range(n):
        lw, lh = int(sizes[2 * l]), int(sizes[2 * l + 1])
        levels.append(out[o:o + lw * lh * 3].reshape(lh, lw, 3))
        o += lw * lh * 3
    return levels


def texture_filter():
    lib = load()
    lib.orc_texture_filter.argtypes = [C.c_void_p]
    out = np.empty(64, np.float32)
    lib.orc_texture_filter(_p(out))
    return out


def texture_lookup(rgb, queries):
    """Anisotropic (EWA) lookups: queries n x 6 (s, t, ds/dx, dt/dx, ds/dy, dt/dy)."""
    lib = load()
    lib.orc_texture_lookup.argtypes = [C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]
    rgb = _f32(rgb)
    q = _f32(queries).reshape(-1, 6)
    out = np.empty((len(q), 3), np.float32)
    lib.orc_texture_lookup(rgb.shape[1], rgb.shape[0], _p(rgb), len(q), _p(q), _p(out))
    return out


def texture_ewa(rgb, level, queries):
    """The EWA filter of one MIP level on its own (queries n x 6 as texture_lookup)."""
    lib = load()
    lib.orc_texture_ewa.argtypes = [C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p,
                                    C.c_void_p]
    rgb = _f32(rgb)
    q = _f32(queries).reshape(-1, 6)
    out = np.empty((len(q), 3), np.float32)
    lib.orc_texture_ewa(rgb.shape[1], rgb.shape[0], _p(rgb), level, len(q), _p(q), _p(out))
    return out


def samples(seed, pixel, n1, n2):
    lib = load()
    lib.orc_samples.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]
    u1 = np.empty(n1, np.float32)
    u2 = np.empty((n2, 2), np.float32)
    lib.orc_samples(seed, pixel, n1, n2, _p(u1), _p(u2))
    return u1, u2


PROBE = {"fresnel_dielectric": (0, 1), "fresnel_conductor": (1, 1), "ggx_d": (2, 1),
         "ggx_sample": (3, 4), "mitchell": (4, 1), "cone_pdf": (5, 1), "coordinate_system": (6, 6),
         "spherical": (7, 2), "dist1d": (8, 4), "roughness_to_alpha": (9, 1)}


def texture_points(s, pfilm):
    """Camera rays through film positions pfilm (2 x n): hit shape and the
    texture-mapped point (6 x n: s, t, ds/dx, dt/dx, ds/dy, dt/dy) a textured
    material reads there."""
    lib = load()
    lib.orc_texture_points.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]
    d, keep = scene_desc(s)
    pf = _f32(pfilm, (2, -1))
    n = pf.shape[1]
    shape = np.empty(n, np.uint32)
    q = np.empty((6, n), np.float32)
    lib.orc_texture_points(C.byref(d), n, _p(pf), _p(shape), _p(q))
    del keep
    return shape, q


def probe(name, *args):
    """One oracle shading function (render.cpp shade_probe) on float32 inputs."""
    lib = load()
    lib.orc_shade_probe.argtypes = [C.c_int, C.c_void_p, C.c_void_p]
    fn, nout = PROBE[name]
    inp = _f32(np.concatenate([np.atleast_1d(np.asarray(a, np.float64)).ravel() for a in args]))
    out = np.zeros(max(nout, 1), np.float32)
    lib.orc_shade_probe(fn, _p(inp), _p(out))
    return out if nout > 1 else float(out[0])


class DiffOracle:
    """Finite-difference reference (diff.cpp): records the base render's
    discrete decisions, then evaluates the L2 image loss in double precision at
    other parameters (alpha per material, eta/k RGB per material, env texels)."""

    def __init__(self, s, nthreads=None):
        lib = load()
        lib.orc_diff_record.argtypes = [C.c_void_p, C.c_int]
        lib.orc_diff_record.restype = C.c_void_p
        lib.orc_diff_loss.argtypes = [C.c_void_p] + [C.c_void_p] * 5 + [C.c_int]
        lib.orc_diff_loss.restype = C.c_double
        lib.orc_diff_free.argtypes = [C.c_void_p]
        self.lib = lib
        self.scene = s
        self.nthreads = nthreads or default_threads()
        d, keep = scene_desc(s)
        self.h = lib.orc_diff_record(C.byref(d), self.nthreads)
        del keep

    def loss(self, alpha, etak, env, target=None, want_rgb=False):
        a = np.ascontiguousarray(alpha, np.float64)
        ek = np.ascontiguousarray(etak, np.float64)
        ev = np.ascontiguousarray(env, np.float64)
        tg = None if target is None else np.ascontiguousarray(target, np.float32)
        rgb = np.empty((self.scene.height, self.scene.width, 3), np.float32) if want_rgb else None
        val = self.lib.orc_diff_loss(self.h, _p(a), _p(ek), _p(ev), _p(tg), _p(rgb), self.nthreads)
        return (val, rgb) if want_rgb else val

    def close(self):
        if self.h:
            self.lib.orc_diff_free(self.h)
            self.h = None

    def __del__(self):
        self.close()


def mesh_bake(xform, positions, indices, normals=None, tangents=None):
    """World-space bake of STriangleMesh::create plus the triangle vertex gather:
    returns (positions_w, normals_w or None, tangents_w or None, tri_verts (nt, 9))."""
    lib = load()
    lib.orc_mesh_bake.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                  C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    xform = _f32(xform).reshape(24)
    pos = _f32(positions).reshape(-1, 3)
    idx = np.ascontiguousarray(indices, np.uint32).reshape(-1, 3)
    assert idx.size == 0 or int(idx.max()) < len(pos)
    nrm = None if normals is None else _f32(normals).reshape(-1, 3)
    tan = None if tangents is None else _f32(tangents).reshape(-1, 3)
    pw = np.empty_like(pos)
    nw = None if nrm is None else np.empty_like(nrm)
    tw = None if tan is None else np.empty_like(tan)
    tri = np.empty((len(idx), 9), np.float32)
    opt = lambda a: None if a is None else _p(a)  # noqa: E731
    lib.orc_mesh_bake(_p(xform), len(pos), _p(pos), opt(nrm), opt(tan), len(idx), _p(idx), _p(pw), opt(nw),
                      opt(tw), _p(tri))
    return pw, nw, tw, tri


def srgb8_to_linear(rgb8):
    """gammaCorrectInv(b / 255.0f) of every byte (SLoadTexture::load's texel decode)."""
    lib = load()
    lib.orc_srgb8_to_linear.argtypes = [C.c_size_t, C.c_void_p, C.c_void_p]
    a = np.ascontiguousarray(rgb8, np.uint8)
    out = np.empty(a.shape, np.float32)
    lib.orc_srgb8_to_linear(a.size, _p(a), _p(out))
    return out
