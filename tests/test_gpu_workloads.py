"""Query parity on every benchmark workload's device-built graph, against the
oracle's heap A* / Dijkstra on the same rows (SVG.cu:241-251, 354-413):

* 64x64 heightfield, K=50: 100 seeded A* queries (one at a time);
* 30K noisy sphere, K=100: 1,000 SSSD queries (shared launches);
* the serving workload on the 72K torus, K=500: 300 requests with Poisson
  arrivals at 1k QPS, 70% A* and 30% MSAD with |S| log-uniform in [1, 64]
  (weights 0), through the continuous-batching server;
* 750K anisotropic terrain, K=100: 1,000 mixed queries (70% SSSD, 30% MSAD).

A* distances within 1e-9 relative, MSAD distances bit-equal, nearest-source
ids those of a shortest-path forest (they may differ from the oracle's only
at exact ties).
"""
import concurrent.futures as cf
import ctypes as C
import math
import os

import numpy as np
import pytest

import oracle as O
import svg_amd as S
from svg_amd import _lib as L
from svg_amd import meshgen as mg
from tests.helpers import assert_source_ids_match, graph_arcs

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def _graph(workload, K):
    xyz, tri = mg.workload(workload)
    m = S.Mesh.from_arrays(xyz, tri)
    g = S.SVG()
    g.AssignMesh(m)
    g.SetParameters(K, 10, 2)
    g.ConstructSVG()
    V, Kg, _ = g.info()
    tails = np.empty(V, np.int32)
    adj = np.empty(V * Kg, np.int32)
    w = np.empty(V * Kg, np.float64)
    S.check(S.lib().svg_graph_to_host(g.g, tails.ctypes.data, adj.ctypes.data, w.ctypes.data,
                                      None, None, None, None))
    og = O.Graph.from_arrays(xyz, adj, w, tails, Kg)
    return m, g, og, (tails, adj, w, Kg)


def _mixed(V, n, seed, msad_frac=0.3):
    """The serving mix: (kind, s, t, sources) with |S| = floor(exp(U[0, ln 65)))."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        if rng.uniform() < msad_frac:
            k = max(1, min(64, int(math.floor(math.exp(rng.uniform(0.0, math.log(65.0)))))))
            out.append((1, 0, 0, rng.choice(V, k, replace=False).astype(np.int32)))
        else:
            s, t = (int(x) for x in rng.choice(V, 2, replace=False))
            out.append((0, s, t, None))
    return out


def _oracle_msad_many(og, queries, want_sid=True):
    """The oracle's Dijkstra for each (sources, weights) on host threads:
    (distances, nearest-source ids)."""
    with cf.ThreadPoolExecutor(THREADS) as ex:
        return list(ex.map(lambda q: og.msad(q[0], q[1])[::2], queries))


def test_heightfield64_astar_100():
    m, g, og, _ = _graph("heightfield64", 50)
    rng = np.random.default_rng(100)
    for i in range(100):
        s, t = (int(x) for x in rng.choice(og.V, 2, replace=False))
        d_gpu = g.SolveSSSD(s, t)
        d_ref, *_ = og.sssd(s, t)
        assert (d_gpu == d_ref) if d_ref == np.finfo(np.float64).max else abs(d_gpu - d_ref) <= 1e-9 * d_ref, \
            (i, s, t, d_gpu, d_ref)


def test_sphere30k_sssd_1000():
    m, g, og, _ = _graph("noisy_sphere", 100)
    rng = np.random.default_rng(1000)
    pairs = rng.integers(0, og.V, size=(1100, 2))
    pairs = pairs[pairs[:, 0] != pairs[:, 1]][:1000]
    print("sphere30k: 1000 SSSD in shared launches", flush=True)
    d_gpu = g.SolveSSSDBatch(pairs)
    d_ref, _ = og.sssd_many(pairs, nthreads=THREADS)
    rel = np.abs(d_gpu - d_ref) / d_ref
    assert np.all(rel <= 1e-9), (int(np.argmax(rel)), float(rel.max()))


def test_serving_mix_on_torus72k_K500():
    """The serving workload itself: the continuous-batching server on the
    benchmark's resident graph (72K torus, K=500), 300 requests of the
    benchmark's mix and arrival process, every answer against the oracle."""
    m, g, og, (tails, adj, w, K) = _graph("bumpy_torus", 500)
    V = og.V
    arcs = graph_arcs(tails, adj, w, K)
    n = 300
    reqs = _mixed(V, n, seed=4)
    rng = np.random.default_rng(44)
    arrival = np.cumsum(rng.exponential(1e-3, n))
    arrival -= arrival[0]
    lib = L.lib()
    srv = C.c_void_p()
    L.check(lib.svg_server_start(g.g, C.byref(L.ServerCfg(64, 0, 64)), C.byref(srv)))
    arr = (L.Request * n)()
    outs, keep = [], []
    for i, (kind, s, t, src) in enumerate(reqs):
        r = arr[i]
        r.kind = kind
        if kind == 0:
            d = np.empty(1)
            r.s, r.t = s, t
            r.dist_out = d.ctypes.data_as(C.POINTER(C.c_double))
            outs.append((d, None))
        else:
            wts = np.zeros(len(src))
            d = np.empty(V)
            sid = np.empty(V, np.int32)
            keep += [src, wts]
            r.src = src.ctypes.data_as(C.POINTER(C.c_int32))
            r.w = wts.ctypes.data_as(C.POINTER(C.c_double))
            r.ns = len(src)
            r.dist_out = d.ctypes.data_as(C.POINTER(C.c_double))
            r.srcid_out = sid.ctypes.data_as(C.POINTER(C.c_int32))
            outs.append((d, sid))
    lat = np.zeros(n)
    print("torus72k: server replay of 300 requests", flush=True)
    L.check(lib.svg_server_replay(srv, arr, np.ascontiguousarray(arrival), n, lat))
    st = L.ServerStats()
    L.check(lib.svg_server_stats_get(srv, C.byref(st)))
    L.check(lib.svg_server_stop(srv))
    assert st.completed == n and np.all(lat >= 0) and st.max_batch_seen > 1
    print("torus72k: oracle answers", flush=True)
    sssd = [(i, s, t) for i, (k, s, t, _) in enumerate(reqs) if k == 0]
    pairs = np.array([[s, t] for _, s, t in sssd], np.int32)
    d_ref, _ = og.sssd_many(pairs, nthreads=THREADS)
    for (i, s, t), dr in zip(sssd, d_ref):
        assert abs(outs[i][0][0] - dr) <= 1e-9 * dr, (i, s, t, outs[i][0][0], dr)
    msad = [(i, src) for i, (k, _, _, src) in enumerate(reqs) if k == 1]
    assert len(msad) > 60
    ref = _oracle_msad_many(og, [(src, np.zeros(len(src))) for _, src in msad], want_sid=True)
    for (i, src), (rd, rsid) in zip(msad, ref):
        d, sid = outs[i]
        assert np.array_equal(d, rd), i
        assert_source_ids_match(arcs, src, np.zeros(len(src)), d, sid, rsid)


def test_terrain750k_mixed_1000():
    """The 750K anisotropic terrain at K=100: 1,000 queries of the 70/30 mix,
    SSSD in shared launches, MSAD through svg_msad_batch in groups of 30."""
    m, g, og, (tails, adj, w, K) = _graph("aniso_terrain", 100)
    V = og.V
    arcs = graph_arcs(tails, adj, w, K)
    del adj, w
    reqs = _mixed(V, 1000, seed=5)
    pairs = np.array([[s, t] for k, s, t, _ in reqs if k == 0], np.int32)
    print(f"terrain750k: {len(pairs)} SSSD", flush=True)
    d_gpu = g.SolveSSSDBatch(pairs)
    d_ref, _ = og.sssd_many(pairs, nthreads=THREADS)
    rel = np.abs(d_gpu - d_ref) / d_ref
    assert np.all(rel <= 1e-9), (int(np.argmax(rel)), float(rel.max()))
    msad = [src for k, _, _, src in reqs if k == 1]
    print(f"terrain750k: {len(msad)} MSAD", flush=True)
    for b in range(0, len(msad), 30):
        group = [(src, np.zeros(len(src))) for src in msad[b:b + 30]]
        dist, sid = g.SolveMSADBatch(group)
        ref = _oracle_msad_many(og, group)
        for j, ((src, wts), (rd, rsid)) in enumerate(zip(group, ref)):
            assert np.array_equal(dist[j], rd), (b + j)
            assert_source_ids_match(arcs, src, wts, dist[j], sid[j], rsid)
        print(f"  msad {b + len(group)} / {len(msad)}", flush=True)
