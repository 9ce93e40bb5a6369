"""GPU surface-point SSSD (SVG::SolveSSSD(f0, p0, f1, p1), SVG.cu:253-352)
against the oracle's restatement on the same graph.

Both ICH runs and the direct answer are the same arithmetic in the same order
as the oracle (bit-equal distances); the graph stage's distances are bit-equal
to the oracle's Dijkstra; the join picks the same vertex.  Parents can differ
only where two graph paths tie exactly (the solver keeps the lowest parent id,
the reference the first relaxation), which these meshes do not produce.
"""
import numpy as np
import pytest

import oracle as O
import svg_amd as S
from svg_amd import meshgen as mg

pytestmark = pytest.mark.gpu

MESHES = {
    "sphere642_K40": (lambda: mg.noisy_sphere(seed=9, splits4=3, centroid_split=False), 40),
    "torus40x24_K80": (lambda: mg.bumpy_torus(nu=40, nv=24), 80),
    "grid12_full": (lambda: mg.planar_grid(n=12, jitter=0.3, seed=7), 144),
    "disk_K40": (lambda: mg.planar_disk(), 40),          # open: boundary faces and vertices
    # K above 511: a single query's heaps and vertex table exceed a block's
    # LDS, so its runs take the global-memory kernel
    "sphere642_K600": (lambda: mg.noisy_sphere(seed=9, splits4=3, centroid_split=False), 600),
}


@pytest.fixture(scope="module", params=sorted(MESHES))
def graph(request):
    make, K = MESHES[request.param]
    xyz, tri = make()
    m = S.Mesh.from_arrays(xyz, tri)
    g = S.SVG()
    g.AssignMesh(m)
    g.SetParameters(K)
    g.ConstructSVG()
    h = g.CopySVGToHost()
    rows = np.zeros(h["adj"].shape, dtype=O.ROW_DTYPE)
    rows["adj"] = h["adj"]
    rows["geodDist"] = h["dist"]
    om = O.Mesh(xyz, tri)
    twin = om.export()["twin"]
    return request.param, xyz, tri, K, g, om, rows, h["tails"], twin


def _same(a, b):
    """Element-wise equality of two result tuples, NaN equal to NaN (a point
    on an edge of its face can give a NaN key-point position on both sides)."""
    return all(x == y or (x != x and y != y) for x, y in zip(a, b))


def _pairs(rng, tri, twin, n):
    out = []
    for i in range(n):
        f0 = int(rng.integers(len(tri)))
        if i % 3 == 2:
            f1 = int(rng.integers(len(tri)))
        else:                               # a short walk: both routes occur
            f1 = f0
            for _ in range(int(rng.integers(0, 6))):
                t = int(twin[3 * f1 + int(rng.integers(3))])
                if t >= 0:
                    f1 = t // 3
        w0, w1 = rng.dirichlet([2, 2, 2]), rng.dirichlet([2, 2, 2])
        out.append((f0, w0, f1, w1))
    return out


def test_point_queries_match_oracle(graph):
    name, xyz, tri, K, g, om, rows, tails, twin = graph
    rng = np.random.default_rng(11)
    routes = {}
    for f0, w0, f1, w1 in _pairs(rng, tri, twin, 36):
        p0, p1 = w0 @ xyz[tri[f0]], w1 @ xyz[tri[f1]]
        print(f"{name} point query f{f0} -> f{f1}", flush=True)
        r = g.SolveSSSDPoints(f0, p0, f1, p1, path=True)
        o = O.sssd_point(om, rows, tails, K, f0, p0, f1, p1)
        routes[o["route"]] = routes.get(o["route"], 0) + 1
        assert r.route == o["route"], (f0, f1, r, o["route"])
        assert r.dist == o["dist"], (f0, f1, r.dist, o["dist"])
        assert (r.src_edge, r.dst_edge, r.last_vert) == (o["src_edge"], o["dst_edge"], o["last_vert"])
        assert r.src_x == o["src_x"] and r.dst_x == o["dst_x"]
        if r.route == 1:
            # the graph path runs from the join vertex back to a seed, and its
            # first vertex's graph distance is its own ICH distance
            assert r.path[0] == r.last_vert
            assert o["gparent"][r.path[-1]] == -1
            assert np.array_equal(r.path[1:], o["gparent"][r.path[:-1]])
        else:
            assert len(r.path) == 0
    assert routes.get(0, 0) > 3 and routes.get(1, 0) > 3, routes


@pytest.mark.parametrize("graph", ["grid12_full"], indirect=True)
def test_point_query_planar_direct_is_euclidean(graph):
    # the closed form holds on the flat mesh only
    name, xyz, tri, K, g, om, rows, tails, twin = graph
    rng = np.random.default_rng(4)
    n = 0
    for f0, w0, f1, w1 in _pairs(rng, tri, twin, 30):
        p0, p1 = w0 @ xyz[tri[f0]], w1 @ xyz[tri[f1]]
        r = g.SolveSSSDPoints(f0, p0, f1, p1)
        if r.route == 0:
            eu = np.linalg.norm(p1 - p0)
            assert abs(r.dist - eu) <= 1e-12 * eu
            n += 1
    assert n > 5


def test_point_query_rejects_bad_faces(graph):
    name, xyz, tri, K, g, *_ = graph
    with pytest.raises(S.SvgError):
        g.SolveSSSDPoints(len(tri), xyz[0], 0, xyz[1])
    with pytest.raises(S.SvgError):
        g.SolveSSSDPoints(0, [np.nan, 0, 0], 1, xyz[1])


def test_point_batch_matches_oracle(graph):
    name, xyz, tri, K, g, om, rows, tails, twin = graph
    rng = np.random.default_rng(23)
    qs = _pairs(rng, tri, twin, 70)
    faces = np.array([(f0, f1) for f0, _, f1, _ in qs], np.int32)
    pts = np.array([(w0 @ xyz[tri[f0]], w1 @ xyz[tri[f1]]) for f0, w0, f1, w1 in qs])
    print(f"{name} point batch of {len(qs)}", flush=True)
    b = g.SolveSSSDPointsBatch(faces, pts)
    for i, (f0, f1) in enumerate(faces):
        o = O.sssd_point(om, rows, tails, K, int(f0), pts[i, 0], int(f1), pts[i, 1])
        got = (b["route"][i], b["dist"][i], b["src_edge"][i], b["dst_edge"][i], b["last_vert"][i],
               b["src_x"][i], b["dst_x"][i])
        want = (o["route"], o["dist"], o["src_edge"], o["dst_edge"], o["last_vert"], o["src_x"], o["dst_x"])
        assert got == want, (i, got, want)
    # the single-query path (LDS-resident runs) agrees with the batch
    for i in range(0, len(qs), 9):
        r = g.SolveSSSDPoints(int(faces[i, 0]), pts[i, 0], int(faces[i, 1]), pts[i, 1])
        assert (r.dist, r.src_edge, r.dst_edge, r.last_vert) == (b["dist"][i], b["src_edge"][i],
                                                                  b["dst_edge"][i], b["last_vert"][i])


def test_point_special_positions_match_oracle(graph):
    """Same face, points on an edge of their face, on boundary faces; a point
    on a vertex stalls both (NaN window keys) and is refused by both."""
    name, xyz, tri, K, g, om, rows, tails, twin = graph
    rng = np.random.default_rng(31)
    boundary = [f for f in range(len(tri)) if (twin[3 * f:3 * f + 3] < 0).any()]
    cases = []
    for i in range(12):
        f0 = int(rng.integers(len(tri)))
        cases.append((f0, rng.dirichlet([2, 2, 2]), f0, rng.dirichlet([2, 2, 2])))        # same face
        f1 = int(rng.integers(len(tri)))
        cases.append((f0, np.array([0.5, 0.5, 0.0]), f1, np.array([0.0, 0.3, 0.7])))     # on edges
        if boundary:
            fb = int(rng.choice(boundary))
            cases.append((fb, rng.dirichlet([2, 2, 2]), f0, rng.dirichlet([2, 2, 2])))
    for f0, w0, f1, w1 in cases:
        p0, p1 = w0 @ xyz[tri[f0]], w1 @ xyz[tri[f1]]
        print(f"{name} special f{f0} {w0.round(2)} -> f{f1} {w1.round(2)}", flush=True)
        r = g.SolveSSSDPoints(f0, p0, f1, p1)
        o = O.sssd_point(om, rows, tails, K, f0, p0, f1, p1)
        got = (r.route, r.dist, r.src_edge, r.dst_edge, r.last_vert, r.src_x, r.dst_x)
        want = (o["route"], o["dist"], o["src_edge"], o["dst_edge"], o["last_vert"], o["src_x"], o["dst_x"])
        assert _same(got, want), (got, want)
    f0, f1 = 0, len(tri) // 2
    p0, p1 = xyz[tri[f0, 0]], xyz[tri[f1, 1]]
    with pytest.raises(RuntimeError):
        O.sssd_point(om, rows, tails, K, f0, p0, f1, p1)
    with pytest.raises(S.SvgError):
        g.SolveSSSDPoints(f0, p0, f1, p1)
    b = g.SolveSSSDPointsBatch(np.array([[f0, f1], [f0, f1]], np.int32),
                               np.array([[p0, p1], [w0 @ xyz[tri[f0]], w1 @ xyz[tri[f1]]]]))
    assert b["route"][0] == -2 and np.isnan(b["dist"][0]) and b["route"][1] >= 0


def test_cli_point_query(tmp_path):
    """svg_cli point (the reference's interactive point driver, kernel.cu:160-202,
    as one command) prints the library's answer."""
    import os
    import subprocess
    xyz, tri = mg.noisy_sphere(seed=9, splits4=3, centroid_split=False)
    obj = tmp_path / "s.obj"
    with open(obj, "w") as fh:
        for p in xyz:
            fh.write("v %.17g %.17g %.17g\n" % tuple(p))
        for t in tri:
            fh.write("f %d %d %d\n" % tuple(t + 1))
    m = S.Mesh()
    assert m.LoadFromFile(str(obj))
    g = S.SVG()
    g.AssignMesh(m)
    g.SetParameters(40)
    g.ConstructSVG()
    f0, f1 = 5, 400
    p0 = np.array([0.2, 0.3, 0.5]) @ xyz[tri[f0]]       # %.17g round-trips exactly
    p1 = np.array([0.6, 0.1, 0.3]) @ xyz[tri[f1]]
    r = g.SolveSSSDPoints(f0, p0, f1, p1)
    cli = os.path.join(os.path.dirname(S.__file__), "svg_cli")
    args = [cli, "point", str(obj), "40", str(f0)] + ["%.17g" % x for x in p0] + [str(f1)] + \
           ["%.17g" % x for x in p1]
    out = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    d = float(out.stdout.split("Distance:")[1].split()[0])
    assert d == r.dist


def test_point_queries_torus72k_K500_match_oracle():
    """The benchmark's graph (bumpy torus, V = 72,000, K = 500): batch and
    single answers equal the oracle's on the device-built rows."""
    xyz, tri = mg.workload("bumpy_torus")
    m = S.Mesh.from_arrays(xyz, tri)
    g = S.SVG()
    g.AssignMesh(m)
    g.SetParameters(500, 10, 2)
    g.ConstructSVG()
    V, K, _ = g.info()
    tails = np.empty(V, np.int32)
    adj = np.empty(V * K, np.int32)
    w = np.empty(V * K, np.float64)
    S.check(S.lib().svg_graph_to_host(g.g, tails.ctypes.data, adj.ctypes.data, w.ctypes.data,
                                      None, None, None, None))
    og = O.Graph.from_arrays(xyz, adj, w, tails, K)
    del adj, w
    om = O.Mesh(xyz, tri)
    twin = om.export()["twin"]
    rng = np.random.default_rng(41)
    qs = _pairs(rng, tri, twin, 64)
    faces = np.array([(f0, f1) for f0, _, f1, _ in qs], np.int32)
    pts = np.array([(w0 @ xyz[tri[f0]], w1 @ xyz[tri[f1]]) for f0, w0, f1, w1 in qs])
    print("torus72k point batch", flush=True)
    b = g.SolveSSSDPointsBatch(faces, pts)
    routes = set()
    for i, (f0, f1) in enumerate(faces):
        o = O.sssd_point(om, og, None, K, int(f0), pts[i, 0], int(f1), pts[i, 1])
        got = (b["route"][i], b["dist"][i], b["src_edge"][i], b["dst_edge"][i], b["last_vert"][i],
               b["src_x"][i], b["dst_x"][i])
        want = (o["route"], o["dist"], o["src_edge"], o["dst_edge"], o["last_vert"], o["src_x"], o["dst_x"])
        assert _same(got, want), (i, got, want)
        routes.add(o["route"])
    assert routes == {0, 1}
    r = g.SolveSSSDPoints(int(faces[0, 0]), pts[0, 0], int(faces[0, 1]), pts[0, 1])
    assert (r.dist, r.route, r.last_vert) == (b["dist"][0], b["route"][0], b["last_vert"][0])


def test_point_batch_spans_chunks():
    """More queries than one chunk of device tables (8,192 at this K): every
    answer still equals the oracle's."""
    xyz, tri = mg.noisy_sphere(seed=9, splits4=3, centroid_split=False)
    K = 40
    m = S.Mesh.from_arrays(xyz, tri)
    g = S.SVG()
    g.AssignMesh(m)
    g.SetParameters(K)
    g.ConstructSVG()
    h = g.CopySVGToHost()
    rows = np.zeros(h["adj"].shape, dtype=O.ROW_DTYPE)
    rows["adj"] = h["adj"]
    rows["geodDist"] = h["dist"]
    og = O.Graph(xyz, rows, h["tails"])
    om = O.Mesh(xyz, tri)
    rng = np.random.default_rng(77)
    n = 8300
    faces = rng.integers(0, len(tri), size=(n, 2)).astype(np.int32)
    w = rng.dirichlet([2, 2, 2], size=(n, 2))
    pts = np.einsum("nsj,nsjk->nsk", w, xyz[tri[faces]])
    print(f"point batch of {n}", flush=True)
    b = g.SolveSSSDPointsBatch(faces, pts)
    for i in list(range(0, n, 151)) + [8191, 8192, n - 1]:
        o = O.sssd_point(om, og, None, K, int(faces[i, 0]), pts[i, 0], int(faces[i, 1]), pts[i, 1])
        got = (b["route"][i], b["dist"][i], b["src_edge"][i], b["dst_edge"][i], b["last_vert"][i])
        want = (o["route"], o["dist"], o["src_edge"], o["dst_edge"], o["last_vert"])
        assert _same(got, want), (i, got, want)


@pytest.mark.parametrize("name", ["sphere", "disk"])
def test_gpu_matches_frozen_oracle_outputs(name):
    """Rows and point queries of the GPU path equal the frozen oracle outputs
    (tests/golden/oracle_small.npz) without running the oracle."""
    import os
    fix = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_small.npz")
    with np.load(fix, allow_pickle=False) as z:
        f = {k[len(name) + 1:]: z[k] for k in z.files if k.startswith(name + "_")}
    K = int(f["K"])
    m = S.Mesh.from_arrays(f["xyz"], f["tri"])
    g = S.SVG()
    g.AssignMesh(m)
    g.SetParameters(K)
    g.ConstructSVG()
    h = g.CopySVGToHost()
    assert np.array_equal(h["tails"], f["tails"])
    for u, t in enumerate(f["tails"]):
        assert np.array_equal(h["adj"][u, :t], f["adj"][u, :t])
        assert np.array_equal(h["dist"][u, :t], f["dist"][u, :t])
        assert np.array_equal(h["src_edge"][u, :t], f["src_edge"][u, :t])
        assert np.array_equal(h["src_x"][u, :t], f["src_x"][u, :t])
    b = g.SolveSSSDPointsBatch(f["faces"], f["pts"])
    got_i = np.stack([b["src_edge"], b["dst_edge"], b["last_vert"], b["route"]], 1)
    got_d = np.stack([b["dist"], b["src_x"], b["dst_x"]], 1)
    assert np.array_equal(got_i, f["point_i"])
    assert np.array_equal(got_d, f["point_d"], equal_nan=True)
