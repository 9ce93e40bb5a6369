"""Serving: continuous batching gives every request the result it gets alone."""
import ctypes as C

import numpy as np
import pytest

import svg_amd as S
from svg_amd import _lib as L
from svg_amd import meshgen as mg

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def graph():
    xyz, tri = mg.bumpy_torus(nu=60, nv=40)
    m = S.Mesh.from_arrays(xyz, tri)
    g = S.SVG()
    g.AssignMesh(m)
    g.SetParameters(60)
    g.ConstructSVG()
    return m, g, len(xyz)


def _replay_matches_alone(g, V, cfg, n=120, seed=3):
    rng = np.random.default_rng(seed)
    lib = L.lib()
    srv = C.c_void_p()
    L.check(lib.svg_server_start(g.g, C.byref(cfg), C.byref(srv)))
    reqs = (L.Request * n)()
    keep, expect = [], []
    for i in range(n):
        r = reqs[i]
        if i % 3 == 2:
            k = int(rng.integers(1, 9))
            src = rng.choice(V, k, replace=False).astype(np.int32)
            w = rng.uniform(0, 0.02, k)
            dout = np.empty(V)
            sout = np.empty(V, np.int32)
            keep += [src, w, dout, sout]
            r.kind = 1
            r.src = src.ctypes.data_as(C.POINTER(C.c_int32))
            r.w = w.ctypes.data_as(C.POINTER(C.c_double))
            r.ns = k
            r.dist_out = dout.ctypes.data_as(C.POINTER(C.c_double))
            r.srcid_out = sout.ctypes.data_as(C.POINTER(C.c_int32))
            d_alone, sid_alone, _ = g.SolveMSAD(src, w)
            expect.append(("msad", dout, sout, d_alone, sid_alone))
        else:
            s, t = (int(x) for x in rng.choice(V, 2, replace=False))
            dout = np.empty(1)
            keep.append(dout)
            r.kind = 0
            r.s, r.t = s, t
            r.dist_out = dout.ctypes.data_as(C.POINTER(C.c_double))
            expect.append(("sssd", dout, None, g.SolveSSSD(s, t), None))
    arrival = np.cumsum(rng.exponential(0.0003, n))
    lat = np.zeros(n)
    L.check(lib.svg_server_replay(srv, reqs, arrival, n, lat))
    st = L.ServerStats()
    L.check(lib.svg_server_stats_get(srv, C.byref(st)))
    L.check(lib.svg_server_stop(srv))
    assert st.completed == n and np.all(lat >= 0)
    for kind, out, sout, ref, sref in expect:
        if kind == "sssd":
            assert out[0] == ref
        else:
            assert np.array_equal(out, ref)
            assert np.array_equal(sout, sref)
    return st


def test_mixed_stream_equals_alone(graph):
    m, g, V = graph
    st = _replay_matches_alone(g, V, L.ServerCfg(16, 0, 8))
    assert st.max_batch_seen > 1, "requests overlapped on the device"


@pytest.mark.parametrize("streams,chain,mds", [(1, -1, 0.0), (1, 64, 8.0), (4, -1, 2.0), (3, 16, 0.5), (8, 4, 1.0)])
def test_server_settings_do_not_change_results(graph, streams, chain, mds):
    # slot groups, chaining budget and MSAD bucket width change the order of
    # relaxations only: every answer stays bit-identical to the query alone
    m, g, V = graph
    cfg = L.ServerCfg(16, 0, 16, 1.0, 4, mds, streams, chain)
    _replay_matches_alone(g, V, cfg, n=60, seed=11 + streams)


def test_queue_full_is_refused_not_lost(graph):
    m, g, V = graph
    lib = L.lib()
    srv = C.c_void_p()
    L.check(lib.svg_server_start(g.g, C.byref(L.ServerCfg(1, 0, 1)), C.byref(srv)))
    outs = []
    tickets = []
    refused = 0
    for i in range(3000):
        d = np.empty(1)
        outs.append(d)
        r = L.Request(0, i % V, (i * 7 + 1) % V, None, None, 0,
                      d.ctypes.data_as(C.POINTER(C.c_double)), None)
        t = lib.svg_server_submit(srv, C.byref(r))
        if t == -6:
            refused += 1
        else:
            assert t >= 0
            tickets.append(t)
    for t in tickets:
        L.check(lib.svg_server_wait(srv, t, None))
    st = L.ServerStats()
    L.check(lib.svg_server_stats_get(srv, C.byref(st)))
    L.check(lib.svg_server_stop(srv))
    assert st.completed == len(tickets)
    assert refused == st.refused
