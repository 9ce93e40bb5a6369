"""ctypes front end of the oracle (CPU restatement of the reference SVG path).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke entry point and by
the benchmark's cpu_baseline leg, always as the checker.  The product path
(svg_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# SVG_ORACLE_LIB names another build of the oracle (a sanitizer build).
LIB_PATH = os.environ.get("SVG_ORACLE_LIB") or os.path.join(HERE, "libsvg_oracle.so")

COUNTERS = [
    "win_push", "win_pop", "win_stale", "pseudo_push", "pseudo_pop", "pseudo_stale",
    "propagate", "child_built", "child_valid", "split_write", "vert_write",
    "fan_halfedge", "pseudo_neigh", "win_drop", "pseudo_drop", "max_winq", "max_pseudoq",
    "reached", "total_calc", "bt_hops", "arcs", "vert_overflow", "split_overflow",
    "vert_keys", "split_keys", "level_skip", "errors", "iterations",
    "winq_left", "pseudoq_left", "total_prev",
]

ROW_DTYPE = np.dtype([
    ("adj", np.int32), ("nextToSrcEdge", np.int32), ("nextToDstEdge", np.int32),
    ("pad", np.int32), ("geodDist", np.float64), ("nextToSrcX", np.float64),
    ("nextToDstX", np.float64),
])

_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_bp = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_rp = np.ctypeslib.ndpointer(dtype=ROW_DTYPE, flags="C_CONTIGUOUS")


def build_lib(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build_lib()
    L = C.CDLL(LIB_PATH)
    L.oracle_mesh_new.restype = C.c_void_p
    L.oracle_mesh_new.argtypes = [_dp, C.c_int, _ip, C.c_int]
    L.oracle_mesh_free.argtypes = [C.c_void_p]
    L.oracle_mesh_export.argtypes = [C.c_void_p, _ip, _dp, _dp, _ip, _bp]
    L.oracle_num_counters.restype = C.c_int
    L.oracle_svg_build.restype = C.c_double
    L.oracle_svg_build.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_uint, C.c_uint, C.c_int,
                                   C.c_int, C.c_int, _rp, _ip, C.c_void_p]
    L.oracle_ich_single.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_uint, C.c_uint, _dp,
                                    _ip, _lp]
    L.oracle_sssd.restype = C.c_double
    L.oracle_sssd.argtypes = [C.c_int, C.c_int, _ip, _dp, _ip, _dp, C.c_int, C.c_int,
                              _dp, _ip, _ip, _lp]
    L.oracle_msad.argtypes = [C.c_int, C.c_int, _ip, _dp, _ip, _dp, _ip, _dp, C.c_int,
                              _dp, _ip, _ip, _lp]
    L.oracle_sssd_many.restype = C.c_double
    L.oracle_sssd_many.argtypes = [C.c_int, C.c_int, _ip, _dp, _ip, _dp, _ip, C.c_int,
                                   C.c_int, _dp]
    L.oracle_ich_point.argtypes = [C.c_void_p, C.c_int, _dp, C.c_int, C.c_uint, C.c_uint, _dp,
                                   _ip, _lp]
    L.oracle_sssd_point.argtypes = [C.c_void_p, C.c_int, C.c_uint, C.c_uint, C.c_int, _ip, _dp,
                                    _ip, C.c_int, _dp, C.c_int, _dp, _dp, _ip, _dp, _ip, _lp]
    L.oracle_selftest_heap.argtypes = [_dp, C.c_int, _ip]
    L.oracle_selftest_hash.argtypes = [C.c_uint, _up, C.c_int, _ip]
    L.oracle_acos_many.argtypes = [_dp, _dp, C.c_int]
    L.oracle_selftest_winq.argtypes = [_dp, C.c_int, _ip, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.oracle_selftest_pseudoq.argtypes = [_dp, C.c_int, _ip, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.oracle_selftest_valid.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.c_double, C.c_int, _ip, _dp, C.c_int]
    L.oracle_selftest_propagate.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double,
                                            C.c_double, C.c_double, C.c_double, C.c_int, _ip, _dp,
                                            C.c_double, C.c_double, _ip, _dp, C.POINTER(C.c_int), _lp,
                                            _dp]
    L.oracle_ich_trace.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _ip, _dp, _dp, _lp]
    assert L.oracle_num_counters() == len(COUNTERS)
    _lib = L
    return L


class Mesh:
    """Half-edge mesh of the oracle (Mesh.cu:24-126 with decisions D1/D2)."""

    def __init__(self, xyz, tri):
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        self.tri = np.ascontiguousarray(tri, dtype=np.int32)
        self.nv = self.xyz.shape[0]
        self.nf = self.tri.shape[0]
        self.h = lib().oracle_mesh_new(self.xyz.ravel(), self.nv, self.tri.ravel(), self.nf)
        if not self.h:
            raise MemoryError("oracle_mesh_new failed")

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.oracle_mesh_free(self.h)
            self.h = None

    def export(self):
        ne = 3 * self.nf
        twin = np.empty(ne, np.int32)
        length = np.empty(ne, np.float64)
        angle = np.empty(self.nv, np.float64)
        first = np.empty(self.nv, np.int32)
        bnd = np.empty(self.nv, np.uint8)
        lib().oracle_mesh_export(self.h, twin, length, angle, first, bnd)
        return dict(twin=twin, len=length, angle=angle, firstEdge=first, boundary=bnd)


def svg_build(mesh: Mesh, K: int, row_cap: int | None = None, s0: int = 0, s1: int | None = None,
              nthreads: int = 0, split_coef: int = 10, vert_coef: int = 2, counters: bool = True):
    """Reference SVG rows for sources [s0, s1): (rows[n, row_cap], tails[n], counters, seconds)."""
    if s1 is None:
        s1 = mesh.nv
    n = s1 - s0
    cap = row_cap if row_cap else K
    rows = np.zeros(n * cap, dtype=ROW_DTYPE)
    tails = np.zeros(n, np.int32)
    cnt = np.zeros(n * len(COUNTERS), np.int64) if counters else None
    secs = lib().oracle_svg_build(mesh.h, K, cap, split_coef, vert_coef, s0, s1, nthreads,
                                  rows, tails,
                                  cnt.ctypes.data_as(C.c_void_p) if counters else None)
    rows = rows.reshape(n, cap)
    if counters:
        cnt = cnt.reshape(n, len(COUNTERS))
    return rows, tails, cnt, secs


def ich_single(mesh: Mesh, src: int, K: int = -1, split_coef: int = 10, vert_coef: int = 2):
    """Distances / enter edges of one ICH run (K = -1 runs to completion)."""
    dist = np.empty(mesh.nv, np.float64)
    enter = np.empty(mesh.nv, np.int32)
    cnt = np.zeros(len(COUNTERS), np.int64)
    lib().oracle_ich_single(mesh.h, src, K, split_coef, vert_coef, dist, enter, cnt)
    return dist, enter, dict(zip(COUNTERS, cnt.tolist()))


def ich_point(mesh: Mesh, face: int, p, K: int = -1, split_coef: int = 10, vert_coef: int = 2):
    """Distances / enter edges of one ICH run from point p of `face`."""
    dist = np.empty(mesh.nv, np.float64)
    enter = np.empty(mesh.nv, np.int32)
    cnt = np.zeros(len(COUNTERS), np.int64)
    pp = np.ascontiguousarray(p, dtype=np.float64).ravel()
    rc = lib().oracle_ich_point(mesh.h, face, pp, K, split_coef, vert_coef, dist, enter, cnt)
    if rc:
        raise ValueError("bad face")
    return dist, enter, dict(zip(COUNTERS, cnt.tolist()))


def sssd_point(mesh: Mesh, rows, tails, K: int, f0: int, p0, f1: int, p1,
               split_coef: int = 10, vert_coef: int = 2):
    """SVG::SolveSSSD(f0, p0, f1, p1) on the graph built with K: `rows` is
    the row array (with `tails`) or a Graph (tails ignored).

    Returns a dict: dist, src_edge, src_x, dst_edge, dst_x, last_vert, route
    (0 direct, 1 through the graph, -1 unreachable), gdist, gparent."""
    V = mesh.nv
    if isinstance(rows, Graph):
        adj, w, tl, Kg = rows.adj, rows.w, rows.tails, rows.K
    else:
        adj = np.ascontiguousarray(rows["adj"], dtype=np.int32).ravel()
        w = np.ascontiguousarray(rows["geodDist"], dtype=np.float64).ravel()
        tl = np.ascontiguousarray(tails, dtype=np.int32)
        Kg = rows.shape[1]
    dout = np.zeros(3, np.float64)
    iout = np.zeros(4, np.int32)
    gdist = np.empty(V, np.float64)
    gpar = np.empty(V, np.int32)
    stat = np.zeros(3, np.int64)
    rc = lib().oracle_sssd_point(mesh.h, K, split_coef, vert_coef, Kg, adj, w, tl,
                                 f0, np.ascontiguousarray(p0, dtype=np.float64).ravel(), f1,
                                 np.ascontiguousarray(p1, dtype=np.float64).ravel(), dout, iout,
                                 gdist, gpar, stat)
    if rc == -2:
        raise RuntimeError("a point run stalled on unordered keys (a point on a vertex of its face)")
    if rc:
        raise ValueError("bad argument")
    return dict(dist=float(dout[0]), src_x=float(dout[1]), dst_x=float(dout[2]),
                src_edge=int(iout[0]), dst_edge=int(iout[1]), last_vert=int(iout[2]),
                route=int(iout[3]), gdist=gdist, gparent=gpar)


class Graph:
    """An SVG in the reference's K-stride row layout (SVG.cu:136)."""

    def __init__(self, xyz, rows, tails):
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float64).ravel()
        self.V = rows.shape[0]
        self.K = rows.shape[1]
        self.adj = np.ascontiguousarray(rows["adj"], dtype=np.int32).ravel()
        self.w = np.ascontiguousarray(rows["geodDist"], dtype=np.float64).ravel()
        self.tails = np.ascontiguousarray(tails, dtype=np.int32)

    @classmethod
    def from_arrays(cls, xyz, adj, w, tails, K: int):
        """From flat K-stride arrays (adj[V*K], w[V*K], tails[V])."""
        g = cls.__new__(cls)
        g.xyz = np.ascontiguousarray(xyz, dtype=np.float64).ravel()
        g.V = len(tails)
        g.K = K
        g.adj = np.ascontiguousarray(adj, dtype=np.int32).ravel()
        g.w = np.ascontiguousarray(w, dtype=np.float64).ravel()
        g.tails = np.ascontiguousarray(tails, dtype=np.int32)
        return g

    @classmethod
    def from_csr(cls, xyz, row_ptr, adj, w):
        V = len(row_ptr) - 1
        deg = np.diff(row_ptr).astype(np.int32)
        K = max(1, int(deg.max()) if V else 1)
        rows = np.zeros((V, K), dtype=ROW_DTYPE)
        for u in range(V):
            a, b = int(row_ptr[u]), int(row_ptr[u + 1])
            rows["adj"][u, : b - a] = adj[a:b]
            rows["geodDist"][u, : b - a] = w[a:b]
        return cls(xyz, rows, deg)

    def sssd(self, s: int, t: int):
        dist = np.empty(self.V, np.float64)
        par = np.empty(self.V, np.int32)
        sid = np.empty(self.V, np.int32)
        stat = np.zeros(3, np.int64)
        d = lib().oracle_sssd(self.V, self.K, self.adj, self.w, self.tails, self.xyz, s, t,
                              dist, par, sid, stat)
        return d, dist, par, sid, stat

    def msad(self, sources, weights):
        src = np.ascontiguousarray(sources, dtype=np.int32)
        wts = np.ascontiguousarray(weights, dtype=np.float64)
        dist = np.empty(self.V, np.float64)
        par = np.empty(self.V, np.int32)
        sid = np.empty(self.V, np.int32)
        stat = np.zeros(3, np.int64)
        lib().oracle_msad(self.V, self.K, self.adj, self.w, self.tails, self.xyz, src, wts,
                          len(src), dist, par, sid, stat)
        return dist, par, sid, stat

    def sssd_many(self, pairs, nthreads: int = 0):
        st = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).ravel())
        out = np.empty(len(st) // 2, np.float64)
        secs = lib().oracle_sssd_many(self.V, self.K, self.adj, self.w, self.tails, self.xyz,
                                      st, len(out), nthreads, out)
        return out, secs


def heap_order(keys):
    keys = np.ascontiguousarray(keys, dtype=np.float64)
    order = np.empty(len(keys), np.int32)
    lib().oracle_selftest_heap(keys, len(keys), order)
    return order


def hash_slots(size: int, keys):
    keys = np.ascontiguousarray(keys, dtype=np.uint32)
    slots = np.empty(len(keys), np.int32)
    ov = lib().oracle_selftest_hash(size, keys, len(keys), slots)
    return slots, bool(ov)


def queue_order(keys, pseudo: bool = False):
    """Push keys into the window (or pseudo-window) queue exactly as the build
    sizes it, then pop everything: (ids in pop order, kept, dropped)."""
    keys = np.ascontiguousarray(keys, dtype=np.float64)
    order = np.full(len(keys), -1, np.int32)
    kept, dropped = C.c_int(), C.c_int()
    f = lib().oracle_selftest_pseudoq if pseudo else lib().oracle_selftest_winq
    f(keys, len(keys), order, C.byref(kept), C.byref(dropped))
    return order[: kept.value], kept.value, dropped.value


def _preset(vdist):
    vids = np.ascontiguousarray(list(vdist.keys()), dtype=np.int32)
    vd = np.ascontiguousarray([vdist[k] for k in vdist], dtype=np.float64)
    return len(vids), vids, vd


def window_valid(mesh: Mesh, edge: int, b0, b1, d0, d1, ps, vdist: dict, is_left: bool) -> bool:
    """IsValidWindow of one window on `edge` with the given vertex distances
    (vertex id -> distance; absent vertices are unreached)."""
    n, vids, vd = _preset(vdist or {0: float(np.finfo(np.float64).max)})
    r = lib().oracle_selftest_valid(mesh.h, edge, b0, b1, d0, d1, ps, n, vids, vd, int(is_left))
    if r < 0:
        raise ValueError("bad edge")
    return bool(r)


def propagate_one(mesh: Mesh, src: int, edge: int, b0, b1, d0, d1, ps, vdist: dict,
                  split=None):
    """PropagateWindow of one window on `edge` from vertex source src, with
    vertex distances preset and the split record of twin(edge) preset to
    split = (dist, x).  Returns dict: children [(edge, b0, b1)] in pop order,
    counters, far_dist, split (dist, x) afterwards."""
    n, vids, vd = _preset(vdist or {src: 0.0})
    ce = np.zeros(2, np.int32)
    cb = np.zeros(4, np.float64)
    nc = C.c_int()
    cnt = np.zeros(len(COUNTERS), np.int64)
    out = np.zeros(3, np.float64)
    big = float(np.finfo(np.float64).max)
    sd, sx = split if split is not None else (big, big)
    r = lib().oracle_selftest_propagate(mesh.h, src, edge, b0, b1, d0, d1, ps, n, vids, vd, sd, sx,
                                        ce, cb, C.byref(nc), cnt, out)
    if r < 0:
        raise ValueError("bad edge")
    kids = [(int(ce[i]), float(cb[2 * i]), float(cb[2 * i + 1])) for i in range(nc.value)]
    return dict(children=kids, counters=dict(zip(COUNTERS, cnt.tolist())), far_dist=float(out[0]),
                split=(float(out[1]), float(out[2])))


def ich_trace(mesh: Mesh, src: int, K: int = -1, cap: int = 100000):
    """Execute(K) from src with its choices recorded: (kind[n] 0 window /
    1 pseudo-window, window-top key[n], pseudo-top key[n], counters)."""
    kind = np.zeros(cap, np.int32)
    wk = np.zeros(cap, np.float64)
    pk = np.zeros(cap, np.float64)
    cnt = np.zeros(len(COUNTERS), np.int64)
    n = lib().oracle_ich_trace(mesh.h, src, K, cap, kind, wk, pk, cnt)
    if n < 0:
        raise ValueError("bad source")
    return kind[:n], wk[:n], pk[:n], dict(zip(COUNTERS, cnt.tolist()))


def acos(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    lib().oracle_acos_many(x, y, len(x))
    return y
