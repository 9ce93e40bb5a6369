"""Shared helpers for the parity tests."""
import numpy as np

import oracle as O


def gpu_rows(svg):
    """GPU graph as (tails, rows dict[V, K])."""
    h = svg.CopySVGToHost()
    return h["tails"], h


def compare_rows(gpu, oracle_rows, oracle_tails, sources, K, fields_exact=True):
    """Bitwise comparison of GPU rows with oracle rows for the given sources.

    Returns a list of mismatch descriptions (empty when identical).
    """
    tails, h = gpu
    bad = []
    for k, s in enumerate(sources):
        t = int(oracle_tails[k])
        if int(tails[s]) != t:
            bad.append(f"source {s}: tail gpu {tails[s]} oracle {t}")
            continue
        o = oracle_rows[k, :t]
        if not np.array_equal(h["adj"][s, :t], o["adj"]):
            bad.append(f"source {s}: neighbour ids differ")
            continue
        if not np.array_equal(h["dist"][s, :t], o["geodDist"]):
            d = np.max(np.abs(h["dist"][s, :t] - o["geodDist"]) / np.maximum(o["geodDist"], 1e-300))
            bad.append(f"source {s}: distances differ (max rel {d:.3g})")
            continue
        if fields_exact:
            for gk, ok in (("src_edge", "nextToSrcEdge"), ("dst_edge", "nextToDstEdge")):
                if not np.array_equal(h[gk][s, :t], o[ok]):
                    bad.append(f"source {s}: {gk} differs")
            for gk, ok, ek in (("src_x", "nextToSrcX", "nextToSrcEdge"),
                               ("dst_x", "nextToDstX", "nextToDstEdge")):
                m = o[ek] != -1
                if not np.array_equal(h[gk][s, :t][m], o[ok][m]):
                    bad.append(f"source {s}: {gk} differs")
    return bad


def oracle_build(xyz, tri, K, sources=None, nthreads=0):
    m = O.Mesh(xyz, tri)
    if sources is None:
        rows, tails, cnt, secs = O.svg_build(m, K, nthreads=nthreads)
        return m, list(range(len(xyz))), rows, tails, cnt
    rows_l, tails_l, cnt_l = [], [], []
    for s in sources:
        r, t, c, _ = O.svg_build(m, K, s0=int(s), s1=int(s) + 1, nthreads=1)
        rows_l.append(r[0]); tails_l.append(t[0]); cnt_l.append(c[0])
    return m, list(sources), np.stack(rows_l), np.array(tails_l), np.stack(cnt_l)


def graph_arcs(tails, adj, w, K):
    """Every arc (u, v, weight) of a K-stride graph, in row order."""
    tails = np.asarray(tails)
    V = len(tails)
    adj = np.asarray(adj).reshape(V, K)
    w = np.asarray(w).reshape(V, K)
    mask = np.arange(K)[None, :] < tails[:, None]
    u = np.repeat(np.arange(V, dtype=np.int64), tails)
    return u, adj[mask].astype(np.int64), w[mask]


def assert_valid_source_ids(arcs, sources, weights, dist, sid, only=None):
    """Nearest-source ids of an MSAD answer are those of a shortest-path
    forest: a source whose distance is its own weight carries its own id;
    every other reached vertex has an arc from a vertex with the same id whose
    distance plus the arc weight is its distance, bit for bit.  Weights are
    positive, so following such arcs strictly decreases the distance and ends
    at a source: the ids are the sources the exact distances come from, and
    two answers can differ only where two sources tie exactly.  `only`
    (a vertex mask) limits the arc test to those vertices: with ids equal to
    a forest's elsewhere, checking where they differ proves the whole set
    (the same induction on distance)."""
    u, v, w = arcs
    if only is not None:
        keep = only[v]
        u, v, w = u[keep], v[keep], w[keep]
    big = np.finfo(np.float64).max
    reached = dist != big
    assert np.all(sid[~reached] == -1)
    own = np.zeros(len(dist), bool)
    for s, ws in zip(np.asarray(sources), np.asarray(weights, dtype=np.float64)):
        if dist[s] == ws:
            own[s] = True
    assert np.all(sid[own] == np.nonzero(own)[0])
    du = dist[u]
    ok = (du != big) & (du + w == dist[v]) & (sid[u] == sid[v])
    has = np.zeros(len(dist), bool)
    has[v[ok]] = True
    need = reached & ~own
    if only is not None:
        need &= only
    assert np.all(has[need]), f"{int(np.sum(need & ~has))} ids without an exact arc from the same source"


def assert_source_ids_match(arcs, sources, weights, dist, sid, ref_sid):
    """sid equals a reference forest's ids except where an exact tie explains
    the difference (see assert_valid_source_ids)."""
    diff = sid != ref_sid
    assert_valid_source_ids(arcs, sources, weights, dist, sid, only=diff)


def assert_valid_parents(arcs, dist, parent):
    """Every parent of a reached non-seed vertex is an exact predecessor."""
    u, v, w = arcs
    big = np.finfo(np.float64).max
    V = len(dist)
    p = np.asarray(parent)
    sel = (dist != big) & (p >= 0)
    keys = u * V + v
    exact = keys[(dist[u] != big) & (dist[u] + w == dist[v])]
    q = np.nonzero(sel)[0]
    assert np.all(np.isin(p[q].astype(np.int64) * V + q, exact)), "a parent is not an exact predecessor"
