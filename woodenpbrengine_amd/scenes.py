"""Seeded synthetic scenes and ray sets for the benchmark and test workloads.

Every array is generated deterministically from a seed with numpy's PCG64
generator (seed = 20261203 + workload index for the named workloads), computed
in float64 and rounded once to float32.  The same arrays feed the GPU path and
the CPU restatement, so parity never depends on the generator.

Shape order everywhere is spheres first, then triangles (the order in which the
reference's ECS iterates CBounds/CCentroid, see include/wpbr.h).
"""
import numpy as np

SEED_BASE = 20261203


def rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


# ---------------------------------------------------------------------------
# transforms
def translate_xform(c):
    """World and inverse 3x4 row-major matrices of a translation, as 24 floats."""
    w = np.zeros((3, 4))
    w[:, :3] = np.eye(3)
    w[:, 3] = c
    inv = np.zeros((3, 4))
    inv[:, :3] = np.eye(3)
    inv[:, 3] = -np.asarray(c, dtype=np.float64)
    return np.concatenate([w.reshape(-1), inv.reshape(-1)]).astype(np.float32)


def random_rotations(g, n):
    """n uniform random rotation matrices (from uniform unit quaternions)."""
    q = g.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


# ---------------------------------------------------------------------------
# LBVH build + traversal microbenchmark (1M triangles, 4M rays)
def random_triangles(n, seed, median_edge=2e-3, sigma=0.5):
    """Equilateral triangles: centroid U[0,1]^3, edge log-normal, random rotation."""
    g = rng(seed)
    c = g.random((n, 3))
    L = np.exp(np.log(median_edge) + sigma * g.standard_normal(n))
    R = random_rotations(g, n)
    ang = np.deg2rad(90.0 + 120.0 * np.arange(3))
    circ = L / np.sqrt(3.0)
    local = np.stack([np.cos(ang), np.sin(ang), np.zeros(3)], axis=1)  # 3 x 3
    v = np.einsum("nij,kj->nki", R, local) * circ[:, None, None] + c[:, None, :]
    return v.reshape(n, 9).astype(np.float32)


def pinhole_rays(width, height, origin, fov_deg=45.0):
    """Primary rays in scanline order, camera looking +z; SoA (6, W*H)."""
    t = np.tan(np.deg2rad(fov_deg) / 2)
    aspect = width / height
    x = (2 * (np.arange(width) + 0.5) / width - 1) * t * aspect
    y = (1 - 2 * (np.arange(height) + 0.5) / height) * t
    X, Y = np.meshgrid(x, y)
    d = np.stack([X.ravel(), Y.ravel(), np.ones(X.size)], axis=0)
    d /= np.linalg.norm(d, axis=0, keepdims=True)
    o = np.repeat(np.asarray(origin, np.float64)[:, None], d.shape[1], axis=1)
    return np.concatenate([o, d], axis=0).astype(np.float32)


def random_rays(n, seed):
    """Origin U[0,1]^3, direction uniform on the sphere; SoA (6, n)."""
    g = rng(seed)
    o = g.random((3, n))
    d = g.standard_normal((3, n))
    d /= np.linalg.norm(d, axis=0, keepdims=True)
    return np.concatenate([o, d], axis=0).astype(np.float32)


def microbench(n_tris=1 << 20, n_primary=(2048, 1024), n_random=1 << 21, seed=SEED_BASE + 1):
    """The LBVH build + traversal microbenchmark: triangles and 2 ray sets."""
    tris = random_triangles(n_tris, seed)
    prim = pinhole_rays(n_primary[0], n_primary[1], (0.5, 0.5, -1.0))
    rnd = random_rays(n_random, seed + 1000)
    rays = np.concatenate([prim, rnd], axis=1)
    return tris, rays


# ---------------------------------------------------------------------------
# geodesic icosphere
def _icosahedron():
    p = (1 + 5 ** 0.5) / 2
    v = np.array([[-1, p, 0], [1, p, 0], [-1, -p, 0], [1, -p, 0], [0, -1, p], [0, 1, p],
                  [0, -1, -p], [0, 1, -p], [p, 0, -1], [p, 0, 1], [-p, 0, -1], [-p, 0, 1]],
                 dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9],
                  [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2],
                  [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
                  [8, 6, 7], [9, 8, 1]])
    return v, f


def icosphere(freq, radius=1.0, center=(0.0, 0.0, 0.0), displace=None):
    """Frequency-`freq` geodesic icosphere: 20*freq^2 triangles (n x 9 float32).

    Points shared by neighbouring faces are computed from the same canonical
    weighted sum (corner indices in ascending order), so the mesh is watertight.
    `displace(unit_dirs) -> radius multipliers` applies a radial displacement.
    """
    V, F = _icosahedron()
    f = int(freq)
    ii, jj = np.meshgrid(np.arange(f + 1), np.arange(f + 1), indexing="ij")
    m = (ii + jj) <= f
    I, J = ii[m], jj[m]
    K = f - I - J
    idx = -np.ones((f + 1, f + 1), dtype=np.int64)
    idx[I, J] = np.arange(len(I))
    # triangles of one face in (i, j) lattice coordinates
    tri = []
    for i in range(f):
        for j in range(f - i):
            tri.append((idx[i, j], idx[i + 1, j], idx[i, j + 1]))
            if i + j < f - 1:
                tri.append((idx[i + 1, j], idx[i + 1, j + 1], idx[i, j + 1]))
    tri = np.array(tri, dtype=np.int64)
    out = []
    for a, b, c in F:
        corners = np.array([a, b, c])
        # point = (a*K + b*I + c*J) / f, summed in ascending corner-index order
        wc = np.stack([K, I, J], axis=1).astype(np.float64)
        order = np.argsort(corners)
        P = np.zeros((len(I), 3))
        for o in order:
            P += V[corners[o]][None, :] * wc[:, o:o + 1]
        P /= f
        P /= np.linalg.norm(P, axis=1, keepdims=True)
        r = radius * (displace(P) if displace is not None else 1.0)
        P = P * (np.asarray(r)[:, None] if np.ndim(r) else r) + np.asarray(center)[None, :]
        out.append(P[tri].reshape(-1, 9))
    return np.concatenate(out, axis=0).astype(np.float32)


def value_noise(seed, octaves=3, base_freq=4.0):
    """Seeded 3-octave value noise in [-1, 1] (hash lattice, smoothstep trilinear)."""
    g = rng(seed)
    table = g.random(1 << 16) * 2 - 1
    perm = g.permutation(1 << 16)

    def lattice(ix, iy, iz):
        h = perm[(ix * 73856093 ^ iy * 19349663 ^ iz * 83492791) & 0xFFFF]
        return table[h]

    def noise(P):
        total = np.zeros(len(P))
        amp, freq, norm = 1.0, base_freq, 0.0
        for _ in range(octaves):
            X = P * freq + 16.0
            i0 = np.floor(X).astype(np.int64)
            t = X - i0
            s = t * t * (3 - 2 * t)
            acc = np.zeros(len(P))
            for dx in (0, 1):
                wx = s[:, 0] if dx else 1 - s[:, 0]
                for dy in (0, 1):
                    wy = s[:, 1] if dy else 1 - s[:, 1]
                    for dz in (0, 1):
                        wz = s[:, 2] if dz else 1 - s[:, 2]
                        acc += wx * wy * wz * lattice(i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz)
            total += amp * acc
            norm += amp
            amp *= 0.5
            freq *= 2.0
        return total / norm

    return noise


# ---------------------------------------------------------------------------
# rendered scenes
N_SPECTRAL = 64
MAT_CONDUCTOR, MAT_DIELECTRIC, MAT_SPECULAR = 0, 1, 2
CU_ETA = (0.2004, 0.9240, 1.1022)   # copper-like eta (RGB)
CU_K = (3.9129, 2.4528, 2.1421)     # copper-like k (RGB)


def bin_centers():
    return 400.0 + (np.arange(N_SPECTRAL) + 0.5) * (300.0 / N_SPECTRAL)


def blackbody_spectrum(T, scale=1.0):
    """Planck's law at the 64 bin centres, normalised to its peak, times scale."""
    lam = bin_centers() * 1e-9
    c, h, kb = 299792458.0, 6.62606957e-34, 1.3806488e-23
    le = (2 * h * c * c) / (lam ** 5 * (np.exp((h * c) / (lam * kb * T)) - 1))
    lam_max = 2.8977721e-3 / T
    le_max = (2 * h * c * c) / (lam_max ** 5 * (np.exp((h * c) / (lam_max * kb * T)) - 1))
    return (le / le_max * scale).astype(np.float32)


def env_map(width, height, seed):
    """Log-normal RGB texels, mu = 0, sigma = 1 (height x width x 3)."""
    g = rng(seed)
    return np.exp(g.standard_normal((height, width, 3))).astype(np.float32)


def identity_xform():
    return translate_xform((0.0, 0.0, 0.0))


def camera(width, height, eye, target, up=(0.0, 1.0, 0.0), fov_deg=45.0, near=0.1, far=10000.0):
    """(raster->camera 4x4, camera->world 3x4) for a perspective camera.

    Raster space follows the reference's screen->raster transform: translate by
    -pMin, scale by 1/(pMax - pMin), then by (res - 1) (CSCameraProjective.h:35-43).
    The screen window is [-a, a] x [-1, 1] with a = width / height."""
    inv_tan = 1.0 / np.tan(np.deg2rad(fov_deg) / 2)
    persp = np.array([[1, 0, 0, 0], [0, 1, 0, 0],
                      [0, 0, far / (far - near), -far * near / (far - near)], [0, 0, 1, 0]], float)
    cam_to_screen = np.diag([inv_tan, inv_tan, 1.0, 1.0]) @ persp
    a = width / height
    pmin, pmax = np.array([-a, -1.0]), np.array([a, 1.0])
    T = np.eye(4)
    T[0, 3], T[1, 3] = -pmin[0], -pmin[1]
    S1 = np.diag([1 / (pmax[0] - pmin[0]), 1 / (pmax[1] - pmin[1]), 1, 1])
    S2 = np.diag([width - 1.0, height - 1.0, 1, 1])
    screen_to_raster = S2 @ S1 @ T
    raster_to_camera = np.linalg.inv(cam_to_screen) @ np.linalg.inv(screen_to_raster)
    eye, target, up = (np.asarray(v, float) for v in (eye, target, up))
    fwd = target - eye
    fwd /= np.linalg.norm(fwd)
    right = np.cross(up / np.linalg.norm(up), fwd)
    right /= np.linalg.norm(right)
    nup = np.cross(fwd, right)
    cam_to_world = np.stack([right, nup, fwd, eye], axis=1)  # 3 x 4
    return raster_to_camera.astype(np.float32).reshape(-1), cam_to_world.astype(np.float32).reshape(-1)


def ground_quad(width, depth, y):
    """createQuad (MeshTools.h:101-155) placed at height y: 2 triangles with the
    quad's uvs, +y normals and +x tangents."""
    w, d = width / 2.0, depth / 2.0
    P = np.array([[-w, y, d], [w, y, d], [-w, y, -d], [w, y, -d]], float)
    uv = np.array([[0, 1], [1, 1], [0, 0], [1, 0]], float)
    idx = np.array([[0, 1, 2], [1, 2, 3]])
    verts = P[idx].reshape(2, 9).astype(np.float32)
    uvs = uv[idx].reshape(2, 6).astype(np.float32)
    normals = np.tile(np.array([0, 1, 0], np.float32), (2, 3))
    tangents = np.tile(np.array([1, 0, 0], np.float32), (2, 3))
    return verts, uvs, normals, tangents


class RenderScene:
    """Arrays describing one rendered scene (spheres first, then triangles)."""

    def __init__(self):
        self.spheres = []        # (center, radius, material or None)
        self.meshes = []         # (verts n x 9, material, uvs, normals, tangents)
        self.materials = []      # 9 floats each
        self.lights = []         # (sphere index, emit spectrum)
        self.env = None          # (rgb h x w x 3, emit spectrum, xform 24)
        self.cam = None
        self.width = self.height = self.spp = 0
        self.seed = 0
        self.sampler = 0         # 0 independent streams, 1 / 2 stratified arrays
        self.textures = []       # RGB textures (h x w x 3 float32, linear)
        self.mat_tex = []        # per material: (metallic, roughness) texture index, -1 none

    def material(self, mtype, alpha=0.1, eta=1.5, eta_rgb=(1.5, 1.5, 1.5), k_rgb=(0, 0, 0), tex_k=-1,
                 tex_r=-1):
        m = np.zeros(9, np.float32)
        m[0] = np.array([mtype], np.uint32).view(np.float32)[0]
        m[1], m[2] = alpha, eta
        m[3:6], m[6:9] = eta_rgb, k_rgb
        self.materials.append(m)
        self.mat_tex.append((tex_k, tex_r))
        return len(self.materials) - 1

    def add_texture(self, rgb):
        self.textures.append(np.ascontiguousarray(rgb, np.float32))
        return len(self.textures) - 1

    def textured_metal(self, tex_k, tex_r, eta=1.5):
        """CMaterialMetal (CMaterialMetal.cpp:24-52): k from the metallic
        texture, alpha from the roughness texture's red value, eta constant."""
        return self.material(MAT_CONDUCTOR, alpha=0.1, eta=eta, eta_rgb=(eta, eta, eta), tex_k=tex_k,
                             tex_r=tex_r)

    def add_sphere(self, center, radius, material=None):
        self.spheres.append((np.asarray(center, float), float(radius), material))
        return len(self.spheres) - 1

    def add_light_sphere(self, center, radius, emit):
        s = self.add_sphere(center, radius, None)
        self.lights.append((s, np.asarray(emit, np.float32)))
        return s

    def add_mesh(self, verts, material, uvs=None, normals=None, tangents=None):
        self.meshes.append((np.asarray(verts, np.float32).reshape(-1, 9), material, uvs, normals,
                            tangents))

    def finalize(self):
        ns = len(self.spheres)
        self.sphere_xform = (np.stack([translate_xform(c) for c, _, _ in self.spheres])
                             if ns else np.zeros((0, 24), np.float32))
        self.sphere_radius = np.array([r for _, r, _ in self.spheres], np.float32)
        verts, flags, uvs, nrm, tan, mats = [], [], [], [], [], []
        for v, m, uv, n, t in self.meshes:
            k = len(v)
            verts.append(v)
            f = (1 if uv is not None else 0) | (2 if (n is not None and t is not None) else 0)
            flags.append(np.full(k, f, np.uint8))
            uvs.append(uv if uv is not None else np.zeros((k, 6), np.float32))
            nrm.append(n if n is not None else np.zeros((k, 9), np.float32))
            tan.append(t if t is not None else np.zeros((k, 9), np.float32))
            mats.append(np.asarray(m, np.uint32) if np.ndim(m) else np.full(k, m, np.uint32))
        self.tri_verts = np.concatenate(verts) if verts else np.zeros((0, 9), np.float32)
        self.tri_flags = np.concatenate(flags) if flags else np.zeros(0, np.uint8)
        self.tri_uv = np.concatenate(uvs).astype(np.float32) if uvs else np.zeros((0, 6), np.float32)
        self.tri_normal = np.concatenate(nrm).astype(np.float32) if nrm else np.zeros((0, 9), np.float32)
        self.tri_tangent = np.concatenate(tan).astype(np.float32) if tan else np.zeros((0, 9), np.float32)
        sph_mat = np.array([0xFFFFFFFF if m is None else m for _, _, m in self.spheres], np.uint32)
        self.shape_material = np.concatenate([sph_mat] + mats).astype(np.uint32)
        self.material_table = (np.stack(self.materials) if self.materials
                               else np.zeros((0, 9), np.float32))
        self.material_textures = np.array(self.mat_tex, np.int32).reshape(-1, 2)
        self.light_shape = np.array([s for s, _ in self.lights], np.uint32)
        self.light_emit = (np.stack([e for _, e in self.lights]).astype(np.float32)
                           if self.lights else np.zeros((0, N_SPECTRAL), np.float32))
        if self.env is not None:
            self.env_rgb, self.env_emit, self.env_xform = self.env
        else:
            self.env_rgb = np.zeros((1, 1, 3), np.float32)
            self.env_emit = np.zeros(N_SPECTRAL, np.float32)
            self.env_xform = identity_xform()
        self.raster_camera, self.camera_world = self.cam
        return self

    @property
    def n_shapes(self):
        return len(self.sphere_radius) + len(self.tri_verts)


def metallic_texture(width, height, seed):
    """A checkerboard of two conductor k colours with noise (linear RGB)."""
    g = rng(seed)
    y, x = np.mgrid[0:height, 0:width]
    cell = ((x * 8 // width) + (y * 4 // height)) % 2
    a, b = np.array([3.6, 2.6, 2.3]), np.array([0.9, 1.4, 3.2])
    rgb = np.where(cell[..., None] == 1, a, b) * (1 + 0.3 * g.random((height, width, 1)))
    return rgb.astype(np.float32)


def roughness_texture(width, height, seed):
    """Smooth-to-rough stripes with noise in the red channel (0 .. 0.6)."""
    g = rng(seed)
    y, x = np.mgrid[0:height, 0:width]
    r = 0.3 + 0.3 * np.sin(2 * np.pi * (3 * x / width + 2 * y / height)) * (0.7 + 0.3 * g.random((height, width)))
    return np.stack([r, 0.5 * r, 0.25 * r], -1).astype(np.float32)


def textured_scene(width=128, height=128, spp=4, seed=SEED_BASE + 5):
    """The plain scene with textured metals (CMaterialMetal: metallic and
    roughness textures, rectangular) on the ground quad (its UVs), on a sphere
    (the sphere mapping) and on the icosphere (each triangle's default UVs)."""
    g = rng(seed)
    s = RenderScene()
    tk = s.add_texture(metallic_texture(256, 128, seed + 1))
    tr = s.add_texture(roughness_texture(64, 96, seed + 2))
    tk2 = s.add_texture(metallic_texture(32, 32, seed + 3))
    ground = s.textured_metal(tk, tr)
    ball = s.textured_metal(tk2, tr, eta=1.2)
    rough_only = s.textured_metal(-1, tr)
    diel = s.material(MAT_DIELECTRIC, alpha=0.1, eta=1.5)
    s.add_sphere((-1.6, -0.5, 0.4), 0.5, ball)
    s.add_sphere((1.6, -0.5, -0.3), 0.5, diel)
    s.add_light_sphere((0.6, 2.0, -1.2), 0.3, blackbody_spectrum(6500.0, 40.0))
    ico = icosphere(8)
    rot = random_rotations(g, 1)[0]
    v = ico.reshape(-1, 3).astype(np.float64)
    v = np.stack([rot[r, 0] * v[:, 0] + rot[r, 1] * v[:, 1] + rot[r, 2] * v[:, 2]
                  for r in range(3)], axis=1)
    s.add_mesh(v.reshape(-1, 9).astype(np.float32), rough_only)
    qv, quv, qn, qt = ground_quad(12.0, 12.0, -1.0)
    s.add_mesh(qv, ground, quv, qn, qt)
    s.materials[rough_only][6:9] = CU_K
    s.env = (env_map(256, 128, seed + 100), np.ones(N_SPECTRAL, np.float32), identity_xform())
    s.cam = camera(width, height, (0.0, 0.9, -4.6), (0.0, -0.2, 0.0))
    s.width, s.height, s.spp, s.seed = width, height, spp, seed
    return s.finalize()


def c0_scene(width=128, height=128, spp=4, seed=SEED_BASE + 0):
    """Plain, CPU-runnable workload: 128x128, 4 spp, direct lighting with MIS."""
    g = rng(seed)
    s = RenderScene()
    cond = s.material(MAT_CONDUCTOR, alpha=0.3, eta_rgb=CU_ETA, k_rgb=CU_K)
    diel = s.material(MAT_DIELECTRIC, alpha=0.1, eta=1.5)
    s.add_sphere((-1.6, -0.5, 0.4), 0.5, diel)
    s.add_sphere((1.6, -0.5, -0.3), 0.5, cond)
    s.add_light_sphere((0.6, 2.0, -1.2), 0.3, blackbody_spectrum(6500.0, 40.0))
    ico = icosphere(32)  # 20 480 triangles, radius 1 at the origin
    rot = random_rotations(g, 1)[0]
    v = ico.reshape(-1, 3).astype(np.float64)
    # element-wise (not BLAS) so shared vertices stay bit-identical
    v = np.stack([rot[r, 0] * v[:, 0] + rot[r, 1] * v[:, 1] + rot[r, 2] * v[:, 2]
                  for r in range(3)], axis=1)
    ico = v.reshape(-1, 9).astype(np.float32)
    s.add_mesh(ico, cond)
    qv, quv, qn, qt = ground_quad(12.0, 12.0, -1.0)
    s.add_mesh(qv, diel, quv, qn, qt)
    s.env = (env_map(256, 128, seed + 100), np.ones(N_SPECTRAL, np.float32), identity_xform())
    s.cam = camera(width, height, (0.0, 0.9, -4.6), (0.0, -0.2, 0.0))
    s.width, s.height, s.spp, s.seed = width, height, spp, seed
    return s.finalize()


def c2_scene(width=1024, height=1024, spp=16, seed=SEED_BASE + 2, freq=317, n_spheres=64,
             env_size=(1024, 512)):
    """Full wavefront render: displaced 2M-triangle icosphere, 64 spheres, ground,
    4 sphere lights and an env map; materials by seed per triangle group."""
    g = rng(seed)
    s = RenderScene()

    def rand_material():
        x = g.random()
        if x < 0.4:
            return s.material(MAT_CONDUCTOR, alpha=float(g.uniform(0.05, 0.6)), eta_rgb=CU_ETA,
                              k_rgb=CU_K)
        if x < 0.8:
            return s.material(MAT_DIELECTRIC, alpha=float(g.uniform(0.01, 0.3)), eta=1.5)
        return s.material(MAT_SPECULAR, eta=1.66)

    # 64 spheres on a ring around the icosphere
    for i in range(n_spheres):
        r = float(g.uniform(0.05, 0.2))
        ang = 2 * np.pi * i / n_spheres + g.uniform(-0.03, 0.03)
        rad = g.uniform(1.6, 3.2)
        c = (rad * np.cos(ang), -1.1 + r + g.uniform(0, 1.5), rad * np.sin(ang))
        s.add_sphere(c, r, rand_material())
    light_pos = [(1.5, 2.5, -1.5), (-2.0, 2.2, -0.5), (0.5, 3.0, 1.5), (-0.8, 2.8, -2.2)]
    for p in light_pos:
        s.add_light_sphere(p, 0.15, blackbody_spectrum(6500.0, 60.0))
    noise = value_noise(seed + 7)
    ico = icosphere(freq, displace=lambda P: 1.0 + 0.05 * noise(P))
    groups = np.array_split(np.arange(len(ico)), 64)
    mat = np.empty(len(ico), np.uint32)
    for grp in groups:
        mat[grp] = rand_material()
    s.add_mesh(ico, mat)
    qv, quv, qn, qt = ground_quad(16.0, 16.0, -1.1)
    s.add_mesh(qv, s.material(MAT_DIELECTRIC, alpha=0.2, eta=1.5), quv, qn, qt)
    s.env = (env_map(env_size[0], env_size[1], seed + 100), np.ones(N_SPECTRAL, np.float32),
             identity_xform())
    s.cam = camera(width, height, (0.0, 1.2, -5.2), (0.0, -0.3, 0.0))
    s.width, s.height, s.spp, s.seed = width, height, spp, seed
    return s.finalize()


def c4_scene(width=512, height=512, spp=8, seed=SEED_BASE + 2, freq=317):
    """Differentiable render: the full render's scene at 512x512, 8 spp, with a
    256x128 env map (its texels are parameters)."""
    return c2_scene(width=width, height=height, spp=spp, seed=seed, freq=freq, env_size=(256, 128))


C3_FILMS = [(128, 128), (256, 256), (512, 512), (1024, 512)]


def c3_scenes(n_scenes=64, spp=8, seed=SEED_BASE + 3, max_triangles=500_000):
    """Batched many-scene render: n_scenes independent scenes of the full
    render's kind (displaced icosphere from a per-scene seed, 64 spheres, ground,
    4 sphere lights, a 256x128 env map, materials as the full render), triangle
    counts log-uniform in [1 000, max_triangles], films drawn from
    128^2 / 256^2 / 512^2 / 1024x512, spp samples per pixel."""
    g = rng(seed)
    out = []
    for k in range(n_scenes):
        tris = float(np.exp(g.uniform(np.log(1e3), np.log(max_triangles))))
        freq = max(2, int(round(np.sqrt(tris / 20.0))))
        w, h = C3_FILMS[int(g.integers(len(C3_FILMS)))]
        out.append(c2_scene(width=w, height=h, spp=spp, seed=seed + 1000 * (k + 1), freq=freq,
                            env_size=(256, 128)))
    return out
