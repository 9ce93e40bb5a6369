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
