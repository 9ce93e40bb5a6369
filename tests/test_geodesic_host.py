"""Arc polylines from the stored key points (host side, no GPU).

The rows come from the oracle's restatement of the reference build.  Every
SVG arc passes no vertex, so its polyline must: start at the source and end at
the target, put each interior point on a mesh edge, cross the last key point's
edge where the row says (nextToDstX), and have the row's geodesic length.
"""
import numpy as np
import pytest

import oracle as O
import svg_amd as S
from svg_amd import meshgen as mg

MESHES = {
    "sphere642": (lambda: mg.noisy_sphere(seed=4, splits4=3, centroid_split=False), 60),
    "torus40x24": (lambda: mg.bumpy_torus(nu=40, nv=24), 80),
    "disk": (lambda: mg.planar_disk(), 40),
}


@pytest.fixture(scope="module", params=sorted(MESHES))
def built(request):
    make, K = MESHES[request.param]
    xyz, tri = make()
    rows, tails, _, _ = O.svg_build(O.Mesh(xyz, tri), K, counters=False)
    return request.param, xyz, tri, rows, tails, S.Mesh.from_arrays(xyz, tri)


def _on_some_edge(p, xyz, tri, tol):
    a = xyz[tri.reshape(-1)]
    b = xyz[np.roll(tri, -1, axis=1).reshape(-1)]
    ab = b - a
    t = np.clip(np.einsum("ij,ij->i", p - a, ab) / np.einsum("ij,ij->i", ab, ab), 0.0, 1.0)
    d = np.linalg.norm(a + t[:, None] * ab - p, axis=1)
    return d.min() <= tol


def test_every_arc_polyline(built):
    name, xyz, tri, rows, tails, mesh = built
    scale = np.ptp(xyz, axis=0).max()
    n_arcs = n_multi = 0
    for u in range(len(xyz)):
        for k in range(tails[u]):
            r = rows[u, k]
            v = int(r["adj"])
            pts, dx = mesh.trace_arc(u, v, int(r["nextToSrcEdge"]), float(r["nextToSrcX"]), int(r["nextToDstEdge"]))
            assert np.array_equal(pts[0], xyz[u]) and np.array_equal(pts[-1], xyz[v])
            length = np.linalg.norm(np.diff(pts, axis=0), axis=1).sum()
            assert abs(length - r["geodDist"]) <= 1e-9 * r["geodDist"], (name, u, v, length, r["geodDist"])
            if r["nextToDstEdge"] >= 0:
                assert abs(dx - r["nextToDstX"]) <= 1e-9 * scale, (name, u, v, dx, r["nextToDstX"])
            n_arcs += 1
            n_multi += len(pts) > 3
    assert n_arcs > 1000 and n_multi > 0
    # interior points lie on mesh edges (a sample of arcs)
    rng = np.random.default_rng(1)
    for u in rng.choice(len(xyz), 20, replace=False):
        for k in range(min(5, tails[u])):
            r = rows[u, k]
            pts, _ = mesh.trace_arc(u, int(r["adj"]), int(r["nextToSrcEdge"]), float(r["nextToSrcX"]),
                                    int(r["nextToDstEdge"]))
            for p in pts[1:-1]:
                assert _on_some_edge(p[None, :], xyz, tri, 1e-12 * scale)


def test_planar_arc_is_a_straight_segment():
    # on a flat jittered grid every geodesic between interior vertices is the
    # Euclidean segment (the grid is convex; its boundary rows are collinear,
    # so arcs along the boundary pass vertices and are left out here)
    xyz, tri = mg.planar_grid(n=12, jitter=0.3, seed=7)
    rows, tails, _, _ = O.svg_build(O.Mesh(xyz, tri), 40, counters=False)
    mesh = S.Mesh.from_arrays(xyz, tri)
    lo, hi = xyz[:, :2].min(axis=0), xyz[:, :2].max(axis=0)
    interior = np.all((xyz[:, :2] > lo + 1e-9) & (xyz[:, :2] < hi - 1e-9), axis=1)
    checked = 0
    for u in np.nonzero(interior)[0][::3]:
        for k in range(tails[u]):
            r = rows[u, k]
            v = int(r["adj"])
            if not interior[v]:
                continue
            pts, _ = mesh.trace_arc(u, v, int(r["nextToSrcEdge"]), float(r["nextToSrcX"]), int(r["nextToDstEdge"]))
            d = xyz[v] - xyz[u]
            # every point is on the segment u-v
            t = (pts - xyz[u]) @ d / (d @ d)
            off = pts - (xyz[u] + t[:, None] * d)
            assert np.abs(off).max() <= 1e-12 and np.all(np.diff(t) >= -1e-15)
            checked += 1
    assert checked > 100


def test_bad_key_point_is_an_error(built):
    name, xyz, tri, rows, tails, mesh = built
    u = 0
    r = rows[u, 0]
    with pytest.raises(S.SvgError):
        mesh.trace_arc(u, int(r["adj"]), 10 ** 9, 0.5, 0)
