"""MI355X Saddle Vertex Graph engine: host-side mirror of the reference API.

``Mesh`` and ``SVG`` keep the reference's method names (Mesh.cuh:32-52,
SVG.cuh:49-82) so callers of the construct-then-solve path switch over
unchanged; every method runs the HIP library through the C ABI
(include/svg_amd.h).
"""
from __future__ import annotations

import ctypes as C
from typing import NamedTuple, Optional

import numpy as np

from ._lib import (SVG_MODE_BATCHED, SVG_MODE_PLAIN, BuildParams, BuildStats, COUNTER_NAMES,
                   SvgError, _ptr, check, lib)

__all__ = ["Mesh", "SVG", "SvgError", "SVG_MODE_PLAIN", "SVG_MODE_BATCHED", "PointResult", "device_name"]


class PointResult(NamedTuple):
    """SVGNode fields of a surface-point query (SVG.cu:253-352)."""
    dist: float
    src_edge: int
    src_x: float
    dst_edge: int
    dst_x: float
    last_vert: int
    route: int
    path: Optional[np.ndarray]


def device_name() -> str:
    buf = C.create_string_buffer(256)
    check(lib().svg_device_name(buf, 256))
    return buf.value.decode()


class Mesh:
    """Half-edge mesh (Mesh.cuh:32-52)."""

    def __init__(self):
        self.h = None

    @classmethod
    def from_arrays(cls, xyz, tri) -> "Mesh":
        m = cls()
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        tri = np.ascontiguousarray(tri, dtype=np.int32)
        h = C.c_void_p()
        check(lib().svg_mesh_from_arrays(xyz.ravel(), xyz.shape[0], tri.ravel(), tri.shape[0],
                                         C.byref(h)))
        m.h = h
        return m

    def LoadFromFile(self, fileName: str) -> bool:            # Mesh.cu:24-126
        self.clear()
        h = C.c_void_p()
        rc = lib().svg_mesh_load_obj(fileName.encode(), C.byref(h))
        if rc < 0:
            return False
        self.h = h
        return True

    def copyToGPU(self) -> bool:                             # Mesh.cu:128-142
        check(lib().svg_mesh_upload(self.h))
        return True

    def _info(self):
        nv, nf, ne = C.c_int(), C.c_int(), C.c_int()
        check(lib().svg_mesh_info(self.h, C.byref(nv), C.byref(nf), C.byref(ne)))
        return nv.value, nf.value, ne.value

    @property
    def vertNum(self) -> int:
        return self._info()[0]

    @property
    def faceNum(self) -> int:
        return self._info()[1]

    @property
    def edgeNum(self) -> int:
        return self._info()[2]

    def trace_arc(self, u, v, src_edge, src_x, dst_edge):
        """Polyline (n, 3) of the SVG arc u -> v from its stored key points, and
        the traced position on the last key point's edge."""
        n = C.c_int()
        dx = C.c_double()
        rc = lib().svg_trace_arc(self.h, int(u), int(v), int(src_edge), float(src_x), int(dst_edge),
                                 np.empty(3, np.float64), 0, C.byref(n), C.byref(dx))
        if rc < 0 and n.value <= 0:
            check(rc)
        out = np.empty((max(n.value, 1), 3), np.float64)
        check(lib().svg_trace_arc(self.h, int(u), int(v), int(src_edge), float(src_x), int(dst_edge),
                                  out.reshape(-1), n.value, C.byref(n), C.byref(dx)))
        return out[: n.value], dx.value

    def export(self):
        nv, nf, ne = self._info()
        twin = np.empty(ne, np.int32)
        ln = np.empty(ne, np.float64)
        ang = np.empty(nv, np.float64)
        first = np.empty(nv, np.int32)
        bnd = np.empty(nv, np.uint8)
        check(lib().svg_mesh_export(self.h, twin, ln, ang, first, bnd))
        return dict(twin=twin, len=ln, angle=ang, firstEdge=first, boundary=bnd)

    def clear(self):                                        # Mesh.cu:144-159
        if self.h:
            lib().svg_mesh_free(self.h)
            self.h = None

    clearGPU = clear

    def __del__(self):
        try:
            self.clear()
        except Exception:
            pass


class SVG:
    """Saddle Vertex Graph (SVG.cuh:49-82)."""

    def __init__(self):
        self.mesh = None
        self.g = None
        self.K = 0
        self.splitInfoCoef = 10
        self.vertInfoCoef = 2

    def AssignMesh(self, mesh: Mesh, d_mesh: Mesh | None = None):     # SVG.cu:109-112
        self.mesh = mesh

    def SetParameters(self, K: int, splitInfoCoef: int = 10, vertInfoCoef: int = 2):
        self.K, self.splitInfoCoef, self.vertInfoCoef = int(K), int(splitInfoCoef), int(vertInfoCoef)

    def Allocation(self) -> bool:
        """Scratch is sized and allocated inside ConstructSVG (SVG.cu:119-140)."""
        return True

    def ConstructSVG(self, mode: int = SVG_MODE_BATCHED, **knobs) -> "SVG":   # SVG.cu:160-182
        if self.mesh is None or self.K <= 0:
            raise SvgError("AssignMesh and SetParameters must precede ConstructSVG")
        self.FreeSVGStructure()
        p = BuildParams()
        lib().svg_build_params_default(C.byref(p), self.K)
        p.split_coef = self.splitInfoCoef
        p.vert_coef = self.vertInfoCoef
        p.mode = mode
        for k, v in knobs.items():
            setattr(p, k, int(v))
        g = C.c_void_p()
        check(lib().svg_build(self.mesh.h, C.byref(p), C.byref(g)))
        self.g = g
        return self

    def Free(self):
        """Build scratch is released by ConstructSVG itself (SVG.cu:142-150)."""

    def FreeSVGStructure(self):                              # SVG.cu:152-158
        if self.g:
            lib().svg_graph_free(self.g)
            self.g = None

    def __del__(self):
        try:
            self.FreeSVGStructure()
        except Exception:
            pass

    # ------------------------------------------------------------------ info
    def info(self):
        V, K, arcs = C.c_int(), C.c_int(), C.c_longlong()
        check(lib().svg_graph_info(self.g, C.byref(V), C.byref(K), C.byref(arcs)))
        return V.value, K.value, arcs.value

    def stats(self) -> dict:
        st = BuildStats()
        check(lib().svg_build_stats_get(self.g, C.byref(st)))
        d = {f: getattr(st, f) for f, _ in BuildStats._fields_ if f != "counters"}
        d["counters"] = dict(zip(COUNTER_NAMES, list(st.counters)))
        return d

    def status(self) -> np.ndarray:
        V, _, _ = self.info()
        out = np.empty(V, np.uint32)
        check(lib().svg_build_status(self.g, out))
        return out

    def CopySVGToHost(self) -> dict:                          # SVG.cu:184-206
        V, K, _ = self.info()
        out = dict(tails=np.empty(V, np.int32), adj=np.empty(V * K, np.int32),
                   dist=np.empty(V * K, np.float64), src_edge=np.empty(V * K, np.int32),
                   dst_edge=np.empty(V * K, np.int32), src_x=np.empty(V * K, np.float64),
                   dst_x=np.empty(V * K, np.float64))
        check(lib().svg_graph_to_host(self.g, *(_ptr(out[k]) for k in
                                                ("tails", "adj", "dist", "src_edge", "dst_edge",
                                                 "src_x", "dst_x"))))
        for k in ("adj", "dist", "src_edge", "dst_edge", "src_x", "dst_x"):
            out[k] = out[k].reshape(V, K)
        return out

    def SaveSVGToFile(self, fileName: str):                   # SVG.cu:208-217
        check(lib().svg_graph_save(self.g, fileName.encode()))

    def LoadSVGFromFile(self, fileName: str):                 # SVG.cu:219-239
        self.FreeSVGStructure()
        g = C.c_void_p()
        check(lib().svg_graph_load(self.mesh.h, self.K, fileName.encode(), C.byref(g)))
        self.g = g

    # --------------------------------------------------------------- queries
    def SolveSSSD(self, *args, path: bool = False):
        """SolveSSSD(s, t): A* between vertices (SVG.cu:241-251); with
        path=True also the vertex path t -> s.  SolveSSSD(f0, p0, f1, p1):
        between surface points (SVG.cu:253-352), see SolveSSSDPoints."""
        if len(args) == 4:
            return self.SolveSSSDPoints(*args, path=path)
        s, t = args
        d = C.c_double()
        if path:
            V, _, _ = self.info()
            buf = np.empty(V, np.int32)
            n = C.c_int()
            check(lib().svg_sssd(self.g, int(s), int(t), C.byref(d), _ptr(buf), V, C.byref(n)))
            return d.value, buf[: n.value].copy()
        check(lib().svg_sssd(self.g, int(s), int(t), C.byref(d), None, 0, None))
        return d.value

    def SolveSSSDPoints(self, f0: int, p0, f1: int, p1, path: bool = False):   # SVG.cu:253-352
        """Geodesic distance between point p0 of face f0 and point p1 of face f1.

        Returns a PointResult: dist, the key points next to each end
        (src_edge, src_x, dst_edge, dst_x: half-edge and distance from its
        verts[0]), last_vert (saddleVertNearDst, -1 when answered directly),
        route (0 direct, 1 through the graph, -1 nothing joins the points) and,
        with path=True, the graph vertices from last_vert back to the first."""
        a = np.ascontiguousarray(p0, dtype=np.float64).reshape(3)
        b = np.ascontiguousarray(p1, dtype=np.float64).reshape(3)
        d = C.c_double()
        ke = np.empty(2, np.int32)
        kx = np.empty(2, np.float64)
        lv = C.c_int32()
        rt = C.c_int()
        n = C.c_int()
        V, _, _ = self.info()
        buf = np.empty(V if path else 1, np.int32)
        check(lib().svg_sssd_point(self.g, int(f0), a, int(f1), b, C.byref(d), ke, kx, C.byref(lv),
                                   C.byref(rt), _ptr(buf), V if path else 0, C.byref(n)))
        return PointResult(d.value, int(ke[0]), float(kx[0]), int(ke[1]), float(kx[1]), lv.value, rt.value,
                           buf[: n.value].copy() if path else None)

    def SolveSSSDPointsBatch(self, faces, points):
        """Many surface-point queries: faces (n, 2) = (f0, f1), points (n, 2, 3)
        = (p0, p1).  Returns dict of arrays: dist, src_edge, src_x, dst_edge,
        dst_x, last_vert, route (as SolveSSSDPoints)."""
        f = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1, 2)
        p = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 2, 3)
        n = f.shape[0]
        if p.shape[0] != n:
            raise ValueError("faces and points differ in length")
        dist = np.empty(n, np.float64)
        ke = np.empty(2 * n, np.int32)
        kx = np.empty(2 * n, np.float64)
        lv = np.empty(n, np.int32)
        rt = np.empty(n, np.int32)
        check(lib().svg_sssd_point_batch(self.g, n, f.reshape(-1), p.reshape(-1), dist, ke, kx, lv, rt))
        return dict(dist=dist, src_edge=ke[0::2].copy(), src_x=kx[0::2].copy(), dst_edge=ke[1::2].copy(),
                    dst_x=kx[1::2].copy(), last_vert=lv, route=rt)

    def SolveMSAD(self, sources, weights, want_parent: bool = True):   # SVG.cu:354-367
        V, _, _ = self.info()
        src = np.ascontiguousarray(sources, dtype=np.int32).ravel()
        w = np.ascontiguousarray(weights, dtype=np.float64).ravel()
        if len(src) == 0 or len(w) != len(src):
            raise ValueError("SolveMSAD needs one weight per source and at least one source")
        if np.any(src < 0) or np.any(src >= V):
            raise ValueError("source vertex out of range")
        dist = np.empty(V, np.float64)
        sid = np.empty(V, np.int32) if want_parent else None
        par = np.empty(V, np.int32) if want_parent else None
        check(lib().svg_msad(self.g, src, w, len(src), dist, _ptr(sid), _ptr(par)))
        return dist, sid, par

    def GeodesicPath(self, verts):
        """Surface polyline (n, 3) of a vertex path of SVG arcs, e.g. the
        reversed path of SolveSSSD(s, t, path=True)."""
        v = np.ascontiguousarray(verts, dtype=np.int32)
        n = C.c_int()
        rc = lib().svg_geodesic_path(self.g, v, len(v), np.empty(3, np.float64), 0, C.byref(n))
        if rc < 0 and n.value <= 0:
            check(rc)
        out = np.empty((max(n.value, 1), 3), np.float64)
        check(lib().svg_geodesic_path(self.g, v, len(v), out.reshape(-1), n.value, C.byref(n)))
        return out[: n.value]

    def SolveMSADBatch(self, queries, want_srcid: bool = True):
        """Many MSAD queries in shared launches: queries is a list of
        (sources, weights).  Returns (dist[nq, V], src_id[nq, V] or None)."""
        V, _, _ = self.info()
        srcs, ws, offs = [], [], [0]
        for s_, w_ in queries:
            s_ = np.asarray(s_, dtype=np.int32).ravel()
            w_ = np.asarray(w_, dtype=np.float64).ravel()
            if len(s_) == 0 or len(s_) != len(w_):
                raise ValueError("every query needs one weight per source and at least one source")
            if np.any(s_ < 0) or np.any(s_ >= V):
                raise ValueError("source vertex out of range")
            srcs.append(s_)
            ws.append(w_)
            offs.append(offs[-1] + len(s_))
        nq = len(srcs)
        src = np.ascontiguousarray(np.concatenate(srcs) if nq else np.zeros(0, np.int32))
        w = np.ascontiguousarray(np.concatenate(ws) if nq else np.zeros(0, np.float64))
        dist = np.empty((max(nq, 1), V), np.float64)
        sid = np.empty((max(nq, 1), V), np.int32) if want_srcid else None
        check(lib().svg_msad_batch(self.g, src, w, np.asarray(offs, np.int64), nq, dist.reshape(-1), _ptr(sid)))
        return dist[:nq], (sid[:nq] if want_srcid else None)

    def SolveSSSDBatch(self, pairs):
        st = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1))
        out = np.empty(len(st) // 2, np.float64)
        check(lib().svg_sssd_batch(self.g, st, len(out), out))
        return out
