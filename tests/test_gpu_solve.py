"""GPU SSSD (A*) / MSAD (Dijkstra) parity with the oracle's heap solvers."""
import numpy as np
import pytest

import oracle as O
import svg_amd as S
from svg_amd import meshgen as mg
from tests.helpers import assert_valid_parents, assert_valid_source_ids, graph_arcs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_graph():
    xyz, tri = mg.noisy_sphere(seed=9, splits4=3, centroid_split=False)
    m = S.Mesh.from_arrays(xyz, tri)
    g = S.SVG()
    g.AssignMesh(m)
    g.SetParameters(40)
    g.ConstructSVG()
    h = g.CopySVGToHost()
    rows = np.zeros(h["adj"].shape, dtype=O.ROW_DTYPE)
    rows["adj"] = h["adj"]
    rows["geodDist"] = h["dist"]
    og = O.Graph(xyz, rows, h["tails"])
    return m, g, og, h


def test_sssd_matches_oracle_astar(small_graph):
    m, g, og, _ = small_graph
    rng = np.random.default_rng(0)
    for _ in range(40):
        s, t = (int(x) for x in rng.choice(og.V, 2, replace=False))
        print(f"sssd {s}->{t}", flush=True)
        d_gpu = g.SolveSSSD(s, t)
        d_ref, *_ = og.sssd(s, t)
        assert abs(d_gpu - d_ref) <= 1e-9 * d_ref, (s, t, d_gpu, d_ref)


def test_sssd_to_an_unreachable_target(small_graph):
    # rows keep the first K qualifying ids (SVG.cu:44-75), so some vertices
    # have no in-arcs: A* to them ends with the oracle's distance (the
    # initial DBL_MAX) and a path that is only the target, s == t gives 0
    m, g, og, h = small_graph
    indeg = np.zeros(og.V, np.int64)
    for u in range(og.V):
        np.add.at(indeg, h["adj"][u, : h["tails"][u]], 1)
    targets = np.nonzero(indeg == 0)[0][:4]
    assert targets.size > 0, "the graph must have a vertex without in-arcs"
    for t in targets:
        s = int((t + 101) % og.V)
        print(f"sssd unreachable {s}->{t}", flush=True)
        d_ref, *_ = og.sssd(s, int(t))
        d_gpu, path = g.SolveSSSD(s, int(t), path=True)
        assert d_gpu == d_ref == np.finfo(np.float64).max
        assert path.tolist() in ([], [int(t)])
    assert g.SolveSSSD(7, 7) == 0.0 == og.sssd(7, 7)[0]


def test_sssd_path_reproduces_distance(small_graph):
    m, g, og, h = small_graph
    d, path = g.SolveSSSD(5, 400, path=True)
    assert path[0] == 400 and path[-1] == 5
    tot = 0.0
    for a, b in zip(path[::-1][:-1], path[::-1][1:]):
        r = h["adj"][a, : h["tails"][a]]
        tot = tot + h["dist"][a, : h["tails"][a]][r == b][0]
    assert tot == d


def test_geodesic_path_polyline(small_graph):
    # the surface polyline of an A* path: from s to t, as long as the distance
    m, g, og, h = small_graph
    xyz = og.xyz.reshape(-1, 3)
    for s, t in ((5, 400), (17, 611), (300, 2)):
        d, path = g.SolveSSSD(s, t, path=True)
        pts = g.GeodesicPath(path[::-1])
        assert np.array_equal(pts[0], xyz[s]) and np.array_equal(pts[-1], xyz[t])
        length = np.linalg.norm(np.diff(pts, axis=0), axis=1).sum()
        assert abs(length - d) <= 1e-9 * d, (s, t, length, d)
        assert len(pts) >= len(path)


@pytest.mark.parametrize("ns", [1, 3, 17])
def test_msad_matches_oracle_dijkstra(small_graph, ns):
    m, g, og, h = small_graph
    rng = np.random.default_rng(ns)
    src = rng.choice(og.V, ns, replace=False).astype(np.int32)
    w = rng.uniform(0, 0.05, ns) if ns > 1 else np.zeros(1)
    print(f"msad ns={ns}", flush=True)
    dist, sid, par = g.SolveMSAD(src, w)
    rdist, rpar, rsid, _ = og.msad(src, w)
    assert np.array_equal(dist, rdist)
    # parents reproduce every distance bit-for-bit and carry the source id;
    # rows keep the first K qualifying ids (SVG.cu:44-75), so some high ids
    # have no in-arcs and stay unreached
    unreached = dist == np.finfo(np.float64).max
    assert np.all(sid[unreached] == -1) and np.all(par[unreached] == -1)
    for v in np.nonzero(~unreached)[0]:
        p = par[v]
        if p == -1:
            assert v in src
            continue
        r = h["adj"][p, : h["tails"][p]]
        wv = h["dist"][p, : h["tails"][p]][r == v][0]
        assert dist[p] + wv == dist[v]
        assert sid[v] == sid[p]
    # both answers are shortest-path forests (ids differ only at exact ties)
    arcs = graph_arcs(h["tails"], h["adj"], h["dist"], h["adj"].shape[1])
    assert_valid_source_ids(arcs, src, w, dist, sid)
    assert_valid_source_ids(arcs, src, w, rdist, rsid)
    assert_valid_parents(arcs, dist, par)
    assert_valid_parents(arcs, rdist, rpar)
    assert np.mean(sid == rsid) > 0.9


@pytest.mark.parametrize("src,w", [([11, 300, 11], [0.05, 0.0, 0.01]),
                                   ([11, 11], [0.0, 0.2]),
                                   ([5, 400, 5, 400], [0.3, 0.1, 0.0, 0.02])])
def test_msad_repeated_sources_keep_their_last_weight(small_graph, src, w):
    # SVG.cu:359-365 writes each source's distance in list order, so a
    # repeated source keeps its last weight, larger or smaller
    m, g, og, h = small_graph
    src = np.array(src, np.int32)
    w = np.array(w)
    print(f"msad repeated {src.tolist()} {w.tolist()}", flush=True)
    dist, sid, par = g.SolveMSAD(src, w)
    rdist, _, rsid, _ = og.msad(src, w)
    assert np.array_equal(dist, rdist)
    last = {int(v): float(x) for v, x in zip(src, w)}
    for v, x in last.items():
        assert dist[v] <= x and sid[v] == rsid[v]
    arcs = graph_arcs(h["tails"], h["adj"], h["dist"], h["adj"].shape[1])
    assert_valid_source_ids(arcs, np.array(list(last)), np.array(list(last.values())), dist, sid)


def test_msad_refuses_weights_outside_the_distance_order(small_graph):
    # distances are ordered by the bits of non-negative doubles: a NaN,
    # infinite or negative weight (and -0.0, whose sign bit is set) is
    # refused before any device work, through every MSAD entry point
    m, g, og, h = small_graph
    src = np.array([3, 200, 500], np.int32)
    for bad in (np.nan, np.inf, -np.inf, -0.5, -0.0):
        w = np.array([0.0, bad, 0.1])
        with pytest.raises(S.SvgError):
            g.SolveMSAD(src, w)
        with pytest.raises(S.SvgError):
            g.SolveMSADBatch([(src, w)])
    dist, _, _ = g.SolveMSAD(src, np.array([0.0, 0.0, 0.1]))
    assert np.array_equal(dist, og.msad(src, [0.0, 0.0, 0.1])[0])


def test_sssd_batch_equals_single(small_graph):
    m, g, og, _ = small_graph
    rng = np.random.default_rng(7)
    pairs = rng.integers(0, og.V, size=(50, 2))
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    out = g.SolveSSSDBatch(pairs)
    for (s, t), d in zip(pairs, out):
        assert d == g.SolveSSSD(int(s), int(t))


def test_svg_file_round_trip_and_layout(small_graph, tmp_path):
    """SaveSVGToFile / LoadSVGFromFile (SVG.cu:208-239): per vertex an int32
    tail, then tail 40-byte SVGNode records (SVG.cuh:21-27: adjNode, 4 bytes
    of padding, geodDist, nextToSrcEdge, nextToDstEdge, nextToSrcX,
    nextToDstX).  The file parses with that layout into the graph's rows, and
    a graph loaded from it answers queries identically."""
    m, g, og, h = small_graph
    path = str(tmp_path / "g.svg")
    g.SaveSVGToFile(path)
    node = np.dtype([("adj", "<i4"), ("pad", "<i4"), ("w", "<f8"), ("se", "<u4"), ("de", "<u4"),
                     ("sx", "<f8"), ("dx", "<f8")])
    assert node.itemsize == 40
    raw = open(path, "rb").read()
    off = 0
    for u in range(len(h["tails"])):
        t = int(np.frombuffer(raw, "<i4", 1, off)[0])
        off += 4
        assert t == h["tails"][u]
        rec = np.frombuffer(raw, node, t, off)
        off += 40 * t
        assert np.array_equal(rec["adj"], h["adj"][u, :t])
        assert np.array_equal(rec["w"], h["dist"][u, :t])
        assert np.array_equal(rec["se"].view(np.int32), h["src_edge"][u, :t])
        assert np.array_equal(rec["de"].view(np.int32), h["dst_edge"][u, :t])
        assert np.array_equal(rec["sx"], h["src_x"][u, :t])
        assert np.array_equal(rec["dx"], h["dst_x"][u, :t])
    assert off == len(raw)
    g2 = S.SVG()
    g2.AssignMesh(m)
    g2.SetParameters(g.K)
    g2.LoadSVGFromFile(path)
    h2 = g2.CopySVGToHost()
    for k in ("tails", "adj", "dist", "src_edge", "dst_edge", "src_x", "dst_x"):
        if k == "tails":
            assert np.array_equal(h2[k], h[k])
        else:
            for u, t in enumerate(h["tails"]):
                assert np.array_equal(h2[k][u, :t], h[k][u, :t]), k
    assert g2.SolveSSSD(5, 400) == g.SolveSSSD(5, 400)
    # a truncated file is an error, not a partial graph
    open(path, "wb").write(raw[: len(raw) // 2])
    g3 = S.SVG()
    g3.AssignMesh(m)
    g3.SetParameters(g.K)
    with pytest.raises(S.SvgError):
        g3.LoadSVGFromFile(path)


def test_cli_build_sssd_msad(small_graph, tmp_path):
    """svg_cli (the reference's driver flows as commands, kernel.cu:17-105,
    288-402, 404-456): build writes the .svg file the library writes, sssd and
    msad print the library's answers."""
    import os
    import subprocess
    m, g, og, h = small_graph
    xyz = og.xyz.reshape(-1, 3)
    tri = np.asarray(mg.noisy_sphere(seed=9, splits4=3, centroid_split=False)[1])
    obj = tmp_path / "s.obj"
    with open(obj, "w") as fh:
        for p in xyz:
            fh.write("v %.17g %.17g %.17g\n" % tuple(p))
        for t in tri:
            fh.write("f %d %d %d\n" % tuple(t + 1))
    cli = os.path.join(os.path.dirname(S.__file__), "svg_cli")
    run = lambda *a: subprocess.run([cli, *a], capture_output=True, text=True, timeout=120)
    out = run("build", str(obj), "40", str(tmp_path / "cli.svg"))
    assert out.returncode == 0, out.stderr
    g.SaveSVGToFile(str(tmp_path / "lib.svg"))
    assert open(tmp_path / "cli.svg", "rb").read() == open(tmp_path / "lib.svg", "rb").read()
    out = run("sssd", str(obj), "40", "5", "400")
    assert out.returncode == 0, out.stderr
    assert float(out.stdout.split("Distance:")[1].split()[0]) == g.SolveSSSD(5, 400)
    out = run("msad", str(obj), "40", "3", "77")
    assert out.returncode == 0, out.stderr
    d, _, _ = g.SolveMSAD([3, 77], [0.0, 0.0])
    far = float(out.stdout.split("vertex at")[1].split()[0])
    assert far == d[d < 1e300].max()
    assert run("nonsense", str(obj), "40").returncode != 0


def test_solvers_torus72k_K500_match_oracle():
    """The benchmark's graph (bumpy torus, V = 72,000, K = 500, 35M arcs):
    A* distances within 1e-9 of the oracle's and MSAD distances bit-equal to
    its heap Dijkstra (SVG.cu:354-413), on the device-built rows."""
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
    arcs = graph_arcs(tails, adj, w, K)
    del adj, w
    rng = np.random.default_rng(72)
    for _ in range(16):
        s, t = (int(x) for x in rng.choice(V, 2, replace=False))
        print(f"torus72k sssd {s}->{t}", flush=True)
        d_gpu = g.SolveSSSD(s, t)
        d_ref, *_ = og.sssd(s, t)
        assert abs(d_gpu - d_ref) <= 1e-9 * d_ref, (s, t, d_gpu, d_ref)
    for ns in (1, 5):
        src = rng.choice(V, ns, replace=False).astype(np.int32)
        wt = rng.uniform(0, 0.05, ns) if ns > 1 else np.zeros(1)
        print(f"torus72k msad ns={ns}", flush=True)
        dist, sid, par = g.SolveMSAD(src, wt)
        rdist, rpar, rsid, _ = og.msad(src, wt)
        assert np.array_equal(dist, rdist)
        assert_valid_source_ids(arcs, src, wt, dist, sid)
        assert_valid_source_ids(arcs, src, wt, rdist, rsid)
        assert_valid_parents(arcs, dist, par)
        assert_valid_parents(arcs, rdist, rpar)
