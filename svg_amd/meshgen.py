"""Seeded synthetic mesh generators for the SVG workloads and the test meshes.

Every generator returns ``(xyz, tri)``: ``xyz`` is float64 ``[V, 3]`` and
``tri`` is int32 ``[F, 3]`` with counter-clockwise corners, exactly what an
OBJ file with ``v x y z`` / ``f a b c`` lines holds after the 1-based ->
0-based shift done by the reference loader (Mesh.cu:41-60).

Randomness comes from a counter-based SplitMix64 stream written here with
numpy integer arithmetic, so the meshes are bit-identical on every machine
that runs the same numpy; ``mesh_digest`` gives the SHA-256 recorded in
results.

Workloads (sizes fixed by the benchmark definitions):

* ``heightfield64``   64x64 grid on the unit square, 8 Gaussian bumps
                      (V=4,096, F=7,938, open).
* ``noisy_sphere``    icosahedron, one 1:3 centroid split, five 1:4 midpoint
                      splits, radial N(0, 0.02) noise (V=30,722, F=61,440).
* ``bumpy_torus``     torus R=1, r=0.35, 360x200 grid, 16 radial Gaussian
                      bumps of amplitude 0.05 (V=72,000, F=144,000).
* ``aniso_terrain``   866x866 grid on [0,4]x[0,1] with 4-octave value noise
                      (V=749,956, F=1,496,450, open).
"""
from __future__ import annotations

import hashlib
import math

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def _mix64(z: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser (Steele, Lea, Flood 2014) on a uint64 array."""
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


class Rng:
    """Counter-based SplitMix64 stream: draw i of seed s is mix(s*G + (i+1)*G)."""

    def __init__(self, seed: int):
        with np.errstate(over="ignore"):
            self._base = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) * _GOLDEN
        self._ctr = 0

    def bits(self, n: int) -> np.ndarray:
        idx = np.arange(self._ctr + 1, self._ctr + 1 + n, dtype=np.uint64)
        self._ctr += n
        with np.errstate(over="ignore"):
            return _mix64(self._base + idx * _GOLDEN)

    def uniform(self, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
        u = (self.bits(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def normal(self, n: int, sigma: float = 1.0) -> np.ndarray:
        u1 = self.uniform(n)
        u2 = self.uniform(n)
        return sigma * np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * math.pi * u2)


def mesh_digest(xyz: np.ndarray, tri: np.ndarray) -> str:
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(xyz, dtype=np.float64).tobytes())
    h.update(np.ascontiguousarray(tri, dtype=np.int32).tobytes())
    return h.hexdigest()


# ----------------------------------------------------------------------------
# grids
# ----------------------------------------------------------------------------

def _grid_tris(nx: int, ny: int, wrap_x: bool = False, wrap_y: bool = False) -> np.ndarray:
    """Quads (i,j)-(i+1,j)-(i+1,j+1)-(i,j+1) split along the (i,j)-(i+1,j+1) diagonal.

    Vertex (i, j) has index j*nx + i.
    """
    qx = nx if wrap_x else nx - 1
    qy = ny if wrap_y else ny - 1
    i, j = np.meshgrid(np.arange(qx), np.arange(qy), indexing="xy")
    i = i.ravel()
    j = j.ravel()
    i1 = (i + 1) % nx
    j1 = (j + 1) % ny
    v00 = j * nx + i
    v10 = j * nx + i1
    v11 = j1 * nx + i1
    v01 = j1 * nx + i
    t0 = np.stack([v00, v10, v11], axis=1)
    t1 = np.stack([v00, v11, v01], axis=1)
    tri = np.empty((2 * t0.shape[0], 3), dtype=np.int32)
    tri[0::2] = t0
    tri[1::2] = t1
    return tri


def heightfield64(seed: int = 1, n: int = 64, bumps: int = 8):
    """n x n heightfield on the unit square with `bumps` seeded Gaussians."""
    rng = Rng(seed)
    amp = rng.uniform(bumps, -0.2, 0.2)
    sig = rng.uniform(bumps, 0.05, 0.2)
    cx = rng.uniform(bumps)
    cy = rng.uniform(bumps)
    g = np.arange(n, dtype=np.float64) / (n - 1)
    x, y = np.meshgrid(g, g, indexing="xy")
    z = np.zeros_like(x)
    for k in range(bumps):
        z += amp[k] * np.exp(-((x - cx[k]) ** 2 + (y - cy[k]) ** 2) / (2.0 * sig[k] ** 2))
    xyz = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    return xyz, _grid_tris(n, n)


def planar_grid(n: int = 12, jitter: float = 0.3, seed: int = 7, scale: float = 1.0):
    """Flat n x n grid on [0,scale]^2 with seeded in-plane jitter of interior vertices.

    Boundary vertices stay on the square, so the domain is convex and the
    geodesic distance between any two vertices is their Euclidean distance.
    """
    rng = Rng(seed)
    h = scale / (n - 1)
    g = np.arange(n, dtype=np.float64) * h
    x, y = np.meshgrid(g, g, indexing="xy")
    x = x.ravel().copy()
    y = y.ravel().copy()
    ii = np.tile(np.arange(n), n)
    jj = np.repeat(np.arange(n), n)
    interior = (ii > 0) & (ii < n - 1) & (jj > 0) & (jj < n - 1)
    m = int(interior.sum())
    x[interior] += rng.uniform(m, -jitter, jitter) * h
    y[interior] += rng.uniform(m, -jitter, jitter) * h
    xyz = np.stack([x, y, np.zeros_like(x)], axis=1)
    return xyz, _grid_tris(n, n)


def planar_disk(n_interior: int = 160, n_boundary: int = 40, seed: int = 13):
    """Flat Delaunay-triangulated disk with a strictly convex boundary.

    Boundary vertices lie on the unit circle at jittered angles, so no three
    are collinear and every geodesic is the straight segment through the
    interior: the distance between any two vertices is their Euclidean one.
    """
    from scipy.spatial import Delaunay

    rng = Rng(seed)
    ang = 2.0 * math.pi * (np.arange(n_boundary) + rng.uniform(n_boundary, -0.3, 0.3)) / n_boundary
    bx, by = np.cos(ang), np.sin(ang)
    rr = 0.9 * np.sqrt(rng.uniform(n_interior))
    th = rng.uniform(n_interior, 0.0, 2.0 * math.pi)
    x = np.concatenate([rr * np.cos(th), bx])
    y = np.concatenate([rr * np.sin(th), by])
    pts = np.stack([x, y], axis=1)
    tri = Delaunay(pts).simplices.astype(np.int32)
    p0, p1, p2 = pts[tri[:, 0]], pts[tri[:, 1]], pts[tri[:, 2]]
    cr = (p1[:, 0] - p0[:, 0]) * (p2[:, 1] - p0[:, 1]) - (p1[:, 1] - p0[:, 1]) * (p2[:, 0] - p0[:, 0])
    flip = cr < 0
    tri[flip] = tri[flip][:, [0, 2, 1]]
    xyz = np.stack([x, y, np.zeros_like(x)], axis=1)
    return xyz, tri


def bumpy_torus(seed: int = 3, nu: int = 360, nv: int = 200, R: float = 1.0,
                r: float = 0.35, bumps: int = 16, amp: float = 0.05):
    """Torus with tube radius r + sum of wrapped Gaussian bumps in (u, v)."""
    rng = Rng(seed)
    cu = rng.uniform(bumps, 0.0, 2.0 * math.pi)
    cv = rng.uniform(bumps, 0.0, 2.0 * math.pi)
    sg = rng.uniform(bumps, 0.1, 0.4)
    u = 2.0 * math.pi * np.arange(nu, dtype=np.float64) / nu
    v = 2.0 * math.pi * np.arange(nv, dtype=np.float64) / nv
    U, V = np.meshgrid(u, v, indexing="xy")
    rho = np.full_like(U, r)
    for k in range(bumps):
        du = np.mod(U - cu[k] + math.pi, 2.0 * math.pi) - math.pi
        dv = np.mod(V - cv[k] + math.pi, 2.0 * math.pi) - math.pi
        rho += amp * np.exp(-(du * du + dv * dv) / (2.0 * sg[k] ** 2))
    x = (R + rho * np.cos(V)) * np.cos(U)
    y = (R + rho * np.cos(V)) * np.sin(U)
    z = rho * np.sin(V)
    xyz = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    return xyz, _grid_tris(nu, nv, wrap_x=True, wrap_y=True)


# ----------------------------------------------------------------------------
# icosphere
# ----------------------------------------------------------------------------

def _icosahedron():
    t = (1.0 + math.sqrt(5.0)) / 2.0
    p = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    p /= np.linalg.norm(p, axis=1, keepdims=True)
    return p, f


def _split4(p: np.ndarray, f: np.ndarray):
    """One 1:4 midpoint split; new vertices are projected to the unit sphere."""
    e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]], axis=0)
    key = np.minimum(e[:, 0], e[:, 1]) * (p.shape[0] + 1) + np.maximum(e[:, 0], e[:, 1])
    uniq, inv = np.unique(key, return_inverse=True)
    a = uniq // (p.shape[0] + 1)
    b = uniq % (p.shape[0] + 1)
    mid = p[a] + p[b]
    mid /= np.linalg.norm(mid, axis=1, keepdims=True)
    base = p.shape[0]
    nf = f.shape[0]
    m01 = base + inv[0:nf]
    m12 = base + inv[nf:2 * nf]
    m20 = base + inv[2 * nf:3 * nf]
    nfaces = np.empty((4 * nf, 3), dtype=np.int64)
    nfaces[0::4] = np.stack([f[:, 0], m01, m20], axis=1)
    nfaces[1::4] = np.stack([m01, f[:, 1], m12], axis=1)
    nfaces[2::4] = np.stack([m20, m12, f[:, 2]], axis=1)
    nfaces[3::4] = np.stack([m01, m12, m20], axis=1)
    return np.concatenate([p, mid], axis=0), nfaces


def _split3(p: np.ndarray, f: np.ndarray):
    """One 1:3 centroid split; centroids are projected to the unit sphere."""
    c = p[f].mean(axis=1)
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    base = p.shape[0]
    ci = base + np.arange(f.shape[0])
    nfaces = np.empty((3 * f.shape[0], 3), dtype=np.int64)
    nfaces[0::3] = np.stack([f[:, 0], f[:, 1], ci], axis=1)
    nfaces[1::3] = np.stack([f[:, 1], f[:, 2], ci], axis=1)
    nfaces[2::3] = np.stack([f[:, 2], f[:, 0], ci], axis=1)
    return np.concatenate([p, c], axis=0), nfaces


def noisy_sphere(seed: int = 2, splits4: int = 5, centroid_split: bool = True,
                 sigma: float = 0.02):
    """Icosphere with radial N(0, sigma) noise; defaults give V=30,722, F=61,440."""
    p, f = _icosahedron()
    if centroid_split:
        p, f = _split3(p, f)
    for _ in range(splits4):
        p, f = _split4(p, f)
    rng = Rng(seed)
    rad = 1.0 + rng.normal(p.shape[0], sigma)
    return p * rad[:, None], f.astype(np.int32)


# ----------------------------------------------------------------------------
# anisotropic value-noise terrain
# ----------------------------------------------------------------------------

def _lattice(seed: int, ix: np.ndarray, iy: np.ndarray) -> np.ndarray:
    """Hash of (seed, ix, iy) mapped to U[-1, 1)."""
    with np.errstate(over="ignore"):
        k = (np.uint64(seed) * np.uint64(0x632BE59BD9B4E019)
             + ix.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)
             + iy.astype(np.uint64) * np.uint64(0xC2B2AE3D27D4EB4F))
    h = _mix64(k)
    return (h >> np.uint64(11)).astype(np.float64) * (2.0 / 9007199254740992.0) - 1.0


def _value_noise(seed: int, x: np.ndarray, y: np.ndarray) -> np.ndarray:
    ix = np.floor(x).astype(np.int64)
    iy = np.floor(y).astype(np.int64)
    fx = x - ix
    fy = y - iy
    sx = fx * fx * fx * (fx * (fx * 6.0 - 15.0) + 10.0)
    sy = fy * fy * fy * (fy * (fy * 6.0 - 15.0) + 10.0)
    n00 = _lattice(seed, ix, iy)
    n10 = _lattice(seed, ix + 1, iy)
    n01 = _lattice(seed, ix, iy + 1)
    n11 = _lattice(seed, ix + 1, iy + 1)
    a = n00 + sx * (n10 - n00)
    b = n01 + sx * (n11 - n01)
    return a + sy * (b - a)


def aniso_terrain(seed: int = 5, n: int = 866, stretch: float = 4.0, octaves: int = 4,
                  base_freq: float = 8.0, amplitude: float = 0.1):
    """n x n grid on [0,stretch]x[0,1]; heights are fractal value noise."""
    g = np.arange(n, dtype=np.float64) / (n - 1)
    U, V = np.meshgrid(g, g, indexing="xy")
    z = np.zeros_like(U)
    norm = 0.0
    amp = 1.0
    freq = base_freq
    for o in range(octaves):
        z += amp * _value_noise(seed * 131 + o, U * freq, V * freq)
        norm += amp
        amp *= 0.5
        freq *= 2.0
    z *= amplitude / norm
    xyz = np.stack([stretch * U.ravel(), V.ravel(), z.ravel()], axis=1)
    return xyz, _grid_tris(n, n)


# ----------------------------------------------------------------------------
# closed-form test meshes
# ----------------------------------------------------------------------------

def tetrahedron():
    """Regular tetrahedron; every vertex has angle sum pi."""
    p = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], dtype=np.float64)
    f = np.array([[0, 1, 2], [0, 3, 1], [0, 2, 3], [1, 3, 2]], dtype=np.int32)
    return p, f


def saddle_cone(n_sectors: int = 10, pts_per_sector: int = 18, wiggle: float = 0.6,
                waves: int = 3, R: float = 4.0, seed: int = 11):
    """Polyhedral cone over a wavy closed spherical polygon, apex = vertex 0.

    Sector k is the planar triangle (apex, R*g_k, R*g_{k+1}) of the cone over
    the polygon g_0..g_{n-1}; it is triangulated on its own (Delaunay in its
    plane) with seeded interior points, so the surface is intrinsically flat
    everywhere except at the apex, whose cone angle theta (the polygon's
    length) exceeds 2*pi: the apex is a saddle.  Interior points avoid the
    sector boundaries, so no vertex lies on a straight spoke from the apex.

    Returns (xyz, tri, info); info["polar"][v] = (rho, alpha) is the
    developed position of vertex v (alpha measured along the polygon) and
    info["theta"] the cone angle.
    """
    from scipy.spatial import Delaunay

    rng = Rng(seed)
    a = 2.0 * math.pi * np.arange(n_sectors) / n_sectors
    g = np.stack([np.cos(a), np.sin(a), wiggle * np.sin(waves * a)], axis=1)
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    gn = np.roll(g, -1, axis=0)
    arc = np.arccos(np.clip(np.sum(g * gn, axis=1), -1.0, 1.0))
    beta = np.concatenate([[0.0], np.cumsum(arc)[:-1]])
    theta = float(arc.sum())
    xyz = [np.zeros(3)]
    polar = [(0.0, 0.0)]
    for k in range(n_sectors):                      # outer corners, vertex 1+k
        xyz.append(R * g[k])
        polar.append((R, beta[k]))
    tris = []
    for k in range(n_sectors):
        k1 = (k + 1) % n_sectors
        u = g[k1] - np.dot(g[k], g[k1]) * g[k]
        u /= np.linalg.norm(u)
        # interior points: alpha strictly inside the sector, rho in (0.6, 0.92 R)
        fr = rng.uniform(pts_per_sector, 0.12, 0.88)
        rr = rng.uniform(pts_per_sector, 0.15, 0.92) * R
        al = fr * arc[k]
        ids = [0, 1 + k, 1 + k1]
        loc = [(0.0, 0.0), (R, 0.0), (R * math.cos(arc[k]), R * math.sin(arc[k]))]
        for j in range(pts_per_sector):
            ids.append(len(xyz))
            xyz.append(rr[j] * (math.cos(al[j]) * g[k] + math.sin(al[j]) * u))
            polar.append((rr[j], beta[k] + al[j]))
            loc.append((rr[j] * math.cos(al[j]), rr[j] * math.sin(al[j])))
        loc = np.asarray(loc)
        dt = Delaunay(loc)
        for s_ in dt.simplices:
            p0, p1, p2 = loc[s_[0]], loc[s_[1]], loc[s_[2]]
            cr = (p1[0] - p0[0]) * (p2[1] - p0[1]) - (p1[1] - p0[1]) * (p2[0] - p0[0])
            t = [ids[s_[0]], ids[s_[1]], ids[s_[2]]]
            if cr < 0:
                t = [t[0], t[2], t[1]]
            tris.append(t)
    info = {"polar": np.asarray(polar), "theta": theta, "beta": beta, "arc": arc}
    return np.asarray(xyz), np.asarray(tris, dtype=np.int32), info


def cone_geodesic(info, a: int, b: int):
    """Closed-form geodesic distance between vertices a and b of ``saddle_cone``.

    With developed angular separation phi (the shorter way round the cone of
    total angle theta), the geodesic is the straight chord when phi < pi and
    runs through the apex (length rho_a + rho_b) otherwise.
    """
    ra, aa = info["polar"][a]
    rb, ab = info["polar"][b]
    d = abs(aa - ab)
    phi = min(d, info["theta"] - d)
    if phi >= math.pi:
        return float(ra + rb), phi
    return float(math.sqrt(ra * ra + rb * rb - 2.0 * ra * rb * math.cos(phi))), phi


WORKLOADS = {
    "heightfield64": dict(fn=heightfield64, K=50, kind="open"),
    "noisy_sphere": dict(fn=noisy_sphere, K=100, kind="closed"),
    "bumpy_torus": dict(fn=bumpy_torus, K=500, kind="closed"),
    "aniso_terrain": dict(fn=aniso_terrain, K=100, kind="open"),
}


def workload(name: str):
    w = WORKLOADS[name]
    return w["fn"]()
