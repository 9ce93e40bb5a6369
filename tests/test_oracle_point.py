"""Surface-point sources in the oracle, pinned by closed forms.

ICH from a point of a face (ICHDevice.cu:583-615 initialisation) must give
the Euclidean distance on a flat convex mesh and the cone closed form (chord,
or through the apex past pi) on a flat saddle cone.  The two-point query
(SVG.cu:253-352) must give the Euclidean distance when it answers directly
through the kept face's windows, min_v |p0 - v| + |v - p1| when it joins
through the graph on a flat mesh whose graph is complete, and key points on
the straight segment.
"""
import math

import numpy as np
import pytest

import oracle as O
from svg_amd import meshgen as mg

BIG = 1e300


def _random_point(rng, xyz, tri, f):
    w = rng.dirichlet([2.0, 2.0, 2.0])
    return w, w @ xyz[tri[f]]


def test_point_ich_planar_euclidean():
    xyz, tri = mg.planar_grid(n=12, jitter=0.3, seed=7)
    m = O.Mesh(xyz, tri)
    rng = np.random.default_rng(0)
    for f in rng.choice(len(tri), 12, replace=False):
        _, p = _random_point(rng, xyz, tri, int(f))
        d, enter, c = O.ich_point(m, int(f), p, -1)
        assert c["errors"] == 0 and c["win_drop"] == 0
        assert np.all(d < BIG)
        eu = np.linalg.norm(xyz - p, axis=1)
        assert np.max(np.abs(d - eu) / eu) < 1e-12
        # the face's own vertices are set directly (no enter edge)
        assert np.all(enter[tri[f]] == -1)


def _cone_polar(info, tri, f, w):
    """Developed polar position of the point with barycentric weights w in f."""
    th, pol = info["theta"], info["polar"]
    vs = tri[f]
    r = np.array([pol[v][0] for v in vs])
    a = np.array([pol[v][1] for v in vs])
    nz = r > 0
    if a[nz].max() - a[nz].min() > th / 2:          # face across alpha = 0
        a = np.where(a < th / 2, a + th, a)
    P = np.stack([r * np.cos(a), r * np.sin(a)], axis=1)
    q = w @ P
    i = int(np.argmax(r))
    off = math.atan2(P[i, 0] * q[1] - P[i, 1] * q[0], P[i, 0] * q[0] + P[i, 1] * q[1])
    return math.hypot(q[0], q[1]), (a[i] + off) % th


def test_point_ich_saddle_cone_closed_form():
    """Chord when the developed angle is below pi, else through the apex.

    Vertices adjacent to the apex are left out for the reason the vertex-source
    cone test gives (only saddle neighbours of a pseudo-source are relaxed)."""
    xyz, tri, info = mg.saddle_cone()
    apex_nb = set(tri[(tri == 0).any(1)].ravel().tolist()) - {0}
    th, pol = info["theta"], info["polar"]
    m = O.Mesh(xyz, tri)
    rng = np.random.default_rng(3)
    n_shadow = checked = 0
    faces = [f for f in rng.permutation(len(tri)) if 0 not in tri[f]][:8]
    for f in faces:
        w, p = _random_point(rng, xyz, tri, int(f))
        rp, ap = _cone_polar(info, tri, int(f), w)
        d, _, c = O.ich_point(m, int(f), p, -1)
        assert c["errors"] == 0 and c["pseudo_push"] > 0
        for v in range(len(xyz)):
            if v == 0 or v in apex_nb:
                continue
            rv, av = pol[v]
            dd = abs(ap - av) % th
            phi = min(dd, th - dd)
            g = rp + rv if phi >= math.pi else math.sqrt(rp * rp + rv * rv - 2 * rp * rv * math.cos(phi))
            n_shadow += phi > math.pi
            # the closed form itself carries the development's rounding
            assert abs(d[v] - g) / g < 1e-11, (f, v, d[v], g)
            checked += 1
    assert n_shadow > 50 and checked > 1000


def test_point_ich_k_stop_on_sphere():
    xyz, tri = mg.noisy_sphere(seed=4, splits4=2, centroid_split=False)
    m = O.Mesh(xyz, tri)
    rng = np.random.default_rng(5)
    for f in [0, 33, 300]:
        _, p = _random_point(rng, xyz, tri, f)
        d, _, c = O.ich_point(m, f, p, 40)
        reached = int((d < BIG).sum())
        # the face's three vertices are set in Initialize without counting
        assert reached == c["total_calc"] + 3
        assert c["total_calc"] >= 40
        eu = np.linalg.norm(xyz - p, axis=1)
        ok = d < BIG
        assert np.all(d[ok] >= eu[ok] * (1 - 1e-12))


def _key_xyz(xyz, tri, e, x):
    f, j = divmod(e, 3)
    a, b = xyz[tri[f, j]], xyz[tri[f, (j + 1) % 3]]
    return a + (b - a) * (x / np.linalg.norm(b - a))


@pytest.fixture(scope="module")
def planar_full():
    xyz, tri = mg.planar_grid(n=12, jitter=0.3, seed=7)
    m = O.Mesh(xyz, tri)
    K = len(xyz)            # every vertex a neighbour: graph distances are Euclidean
    rows, tails, _, _ = O.svg_build(m, K, counters=False)
    return xyz, tri, m, K, rows, tails


def test_sssd_point_planar_direct(planar_full):
    xyz, tri, m, K, rows, tails = planar_full
    rng = np.random.default_rng(0)
    scale = np.ptp(xyz, axis=0).max()
    n_direct = 0
    for _ in range(120):
        f0, f1 = (int(x) for x in rng.integers(0, len(tri), 2))
        _, p0 = _random_point(rng, xyz, tri, f0)
        _, p1 = _random_point(rng, xyz, tri, f1)
        r = O.sssd_point(m, rows, tails, K, f0, p0, f1, p1)
        if r["route"] != 0:
            continue
        n_direct += 1
        eu = np.linalg.norm(p1 - p0)
        assert abs(r["dist"] - eu) <= 1e-12 * eu
        assert r["last_vert"] == -1
        # the destination key point is on f1's boundary, the source one leaves f0
        assert r["dst_edge"] // 3 == f1
        twin = m.export()["twin"]
        assert twin[r["src_edge"]] // 3 == f0
        # both on the segment p0 -> p1, the source one first
        d = p1 - p0
        ts = []
        for e, x in ((r["src_edge"], r["src_x"]), (r["dst_edge"], r["dst_x"])):
            q = _key_xyz(xyz, tri, e, x)
            t = (q - p0) @ d / (d @ d)
            assert np.linalg.norm(p0 + t * d - q) <= 1e-12 * scale
            ts.append(t)
        assert 0.0 <= ts[0] <= ts[1] <= 1.0
    assert n_direct > 80


def test_sssd_point_planar_through_graph(planar_full):
    """Points in the same or adjacent faces join through a vertex."""
    xyz, tri, m, K, rows, tails = planar_full
    rng = np.random.default_rng(1)
    twin = m.export()["twin"]
    n_graph = 0
    for f0 in rng.choice(len(tri), 40, replace=False):
        f0 = int(f0)
        nb = [f0] + [int(twin[3 * f0 + j]) // 3 for j in range(3) if twin[3 * f0 + j] >= 0]
        f1 = int(rng.choice(nb))
        _, p0 = _random_point(rng, xyz, tri, f0)
        _, p1 = _random_point(rng, xyz, tri, f1)
        r = O.sssd_point(m, rows, tails, K, f0, p0, f1, p1)
        if r["route"] != 1:
            continue
        n_graph += 1
        via = np.linalg.norm(xyz - p0, axis=1) + np.linalg.norm(xyz - p1, axis=1)
        assert abs(r["dist"] - via.min()) <= 1e-12 * via.min()
        v = r["last_vert"]
        assert abs(via[v] - r["dist"]) <= 1e-12 * r["dist"]
    assert n_graph > 10


def test_sssd_point_graph_route_composes_pinned_parts():
    """On a curved mesh with K-limited runs the graph route is the multi-source
    Dijkstra seeded by the p0 run joined with the p1 run (each pinned above)."""
    xyz, tri = mg.noisy_sphere(seed=4, splits4=2, centroid_split=False)
    m = O.Mesh(xyz, tri)
    K = 30
    rows, tails, _, _ = O.svg_build(m, K, counters=False)
    g = O.Graph(xyz, rows, tails)
    rng = np.random.default_rng(2)
    twin = m.export()["twin"]
    seen = {0: 0, 1: 0}
    for _ in range(60):
        f0 = int(rng.integers(len(tri)))
        f1 = f0
        for _ in range(int(rng.integers(1, 5))):        # a short walk across edges
            f1 = int(twin[3 * f1 + int(rng.integers(3))]) // 3
        _, p0 = _random_point(rng, xyz, tri, f0)
        _, p1 = _random_point(rng, xyz, tri, f1)
        r = O.sssd_point(m, rows, tails, K, f0, p0, f1, p1)
        seen[r["route"]] += 1
        eu = np.linalg.norm(p1 - p0)
        assert r["dist"] >= eu * (1 - 1e-12)
        if r["route"] != 1:
            continue
        d0, _, _ = O.ich_point(m, f0, p0, K)
        d1, _, _ = O.ich_point(m, f1, p1, K)
        s0 = np.nonzero(d0 < BIG)[0]
        D0, _, _, _ = g.msad(s0, d0[s0])
        r1 = d1 < BIG
        assert np.array_equal(D0, r["gdist"])
        assert r["dist"] == np.min(D0[r1] + d1[r1])
    assert seen[0] > 5 and seen[1] > 5


def test_point_on_a_vertex_stalls_and_edge_points_are_exact(planar_full):
    """A point on a vertex of its face puts the source on a window's line (NaN
    keys, where the reference loops): the query is refused.  A point on an
    edge of its face is an ordinary source: direct answers stay Euclidean to
    1e-8 (the face-point placement, ICHDevice.cu:186-197, takes a zero
    Heron area as the square root of rounding noise)."""
    xyz, tri, m, K, rows, tails = planar_full
    with pytest.raises(RuntimeError):
        O.sssd_point(m, rows, tails, K, 3, xyz[tri[3, 0]], 100, xyz[tri[100, 1]])
    rng = np.random.default_rng(8)
    n = 0
    for _ in range(40):
        f0, f1 = (int(x) for x in rng.integers(0, len(tri), 2))
        p0 = np.array([0.5, 0.5, 0.0]) @ xyz[tri[f0]]
        p1 = np.array([0.0, 0.3, 0.7]) @ xyz[tri[f1]]
        r = O.sssd_point(m, rows, tails, K, f0, p0, f1, p1)
        if r["route"] == 0:
            eu = np.linalg.norm(p1 - p0)
            assert abs(r["dist"] - eu) <= 1e-8 * eu
            n += 1
    assert n > 20
