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
        w = np.stack([I, J, K], axis=1).astype(np.float64)  # weights of b, c, a below
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
        del w
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
