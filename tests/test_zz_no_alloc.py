"""svg_msad_noalloc against svg_msad through the C ABI: distances, nearest-
source ids and parents bit-equal, on two calls in a row (the second with a
smaller graph and fewer sources, so the array holds the first call's larger
contents), and a graph beyond the array's limit refused."""
import ctypes as C

import numpy as np
import pytest

import svg_amd as S
from svg_amd import _lib as L
from svg_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def _graph(xyz, tri, K):
    m = S.Mesh.from_arrays(xyz, tri)
    g = S.SVG()
    g.AssignMesh(m)
    g.SetParameters(K, 10, 2)
    g.ConstructSVG()
    return m, g


def _both(g, V, src, w):
    lib = L.lib()
    out = []
    for fn in (lib.svg_msad, lib.svg_msad_noalloc):
        d = np.empty(V)
        sid = np.empty(V, np.int32)
        par = np.empty(V, np.int32)
        L.check(fn(g.g, src, w, len(src), d, sid.ctypes.data_as(C.c_void_p), par.ctypes.data_as(C.c_void_p)))
        out.append((d, sid, par))
    return out


def test_noalloc_msad_equals_msad_twice():
    rng = np.random.default_rng(7)
    # call 1: the 72K torus at K=500 (the serving graph), 17 sources
    xyz, tri = mg.workload("bumpy_torus")
    m1, g1 = _graph(xyz, tri, 500)
    V1 = len(xyz)
    src = rng.choice(V1, 17, replace=False).astype(np.int32)
    w = rng.uniform(0, 0.05, 17)
    print("noalloc call 1: torus72k K=500, 17 sources", flush=True)
    (d0, s0, p0), (d1, s1, p1) = _both(g1, V1, src, w)
    assert np.array_equal(d0, d1) and np.array_equal(s0, s1) and np.array_equal(p0, p1)
    assert np.sum(d1 < 1e300) > V1 // 2
    # call 2: a smaller graph and input (642 vertices, 3 sources) over the
    # array the first call left filled
    xyz2, tri2 = mg.noisy_sphere(seed=4, splits4=3, centroid_split=False)
    m2, g2 = _graph(xyz2, tri2, 60)
    V2 = len(xyz2)
    src2 = np.array([5, 100, 333], np.int32)
    w2 = np.array([0.0, 0.01, 0.02])
    print("noalloc call 2: sphere642 K=60, 3 sources", flush=True)
    (e0, t0, q0), (e1, t1, q1) = _both(g2, V2, src2, w2)
    assert np.array_equal(e0, e1) and np.array_equal(t0, t1) and np.array_equal(q0, q1)
    # and the larger graph once more, after the small call
    print("noalloc call 3: torus72k again, 1 source", flush=True)
    (d0, s0, p0), (d1, s1, p1) = _both(g1, V1, src[:1], np.zeros(1))
    assert np.array_equal(d0, d1) and np.array_equal(s0, s1) and np.array_equal(p0, p1)


def test_noalloc_refuses_graphs_beyond_its_array():
    # 360 x 300 torus: 108,000 vertices, beyond the 64 MiB array's ~87,000
    xyz, tri = mg.bumpy_torus(nu=360, nv=300)
    m, g = _graph(xyz, tri, 10)
    V = len(xyz)
    d = np.empty(V)
    print("noalloc refusal: 108K vertices", flush=True)
    rc = L.lib().svg_msad_noalloc(g.g, np.array([0], np.int32), np.zeros(1), 1, d, None, None)
    assert rc == -1
    assert b"too large" in L.lib().svg_last_error()
    # bad arguments are refused before any device work
    assert L.lib().svg_msad_noalloc(g.g, np.array([V], np.int32), np.zeros(1), 1, d, None, None) == -1
