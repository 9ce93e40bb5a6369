// Geodesic polylines of SVG arcs and of vertex paths.
//
// An SVG arc u -> v is a geodesic that passes no vertex (SVG.cu:56), so in
// the unfolding of the faces it crosses it is one straight segment.  The row
// keeps the crossing next to each end (SVGNode::nextToSrcEdge/X and
// nextToDstEdge/X, SVG.cuh:21-27; set by BuildGeodesicPathTo,
// ICHDevice.cu:336-480): the segment leaves u through the first and enters
// v's face through the last.  The crossings in between follow by unfolding
// face after face along the straight line, the construction the reference
// sketches for tracing a geodesic from a key point (InitialValueGeodesic.cu:
// 54-256, not called there).  A key point (e, x) is the point at distance x
// from verts[0] of half-edge e; the crossing half-edge belongs to the face on
// v's side of the crossing, as in the backtrace.
#include <cmath>
#include <string>
#include <vector>

#include "svg_mesh.h"

namespace svg {

namespace {

inline int next_he(int e) { return (e % 3 == 2) ? e - 2 : e + 1; }
inline int prev_he(int e) { return (e % 3 == 0) ? e + 2 : e - 1; }
inline int v0_of(const HostMesh& m, int e) { return m.tri[e]; }
inline int v1_of(const HostMesh& m, int e) { return m.tri[3 * (e / 3) + (e % 3 + 1) % 3]; }

void key_point_xyz(const HostMesh& m, int e, double x, double* out)
{
    const int a = v0_of(m, e), b = v1_of(m, e);
    const double f = x / m.len[e];
    for (int k = 0; k < 3; ++k) out[k] = m.xyz[3 * a + k] + (m.xyz[3 * b + k] - m.xyz[3 * a + k]) * f;
}

// Position of a point with distances dA, dB from A = (0, 0) and B = (l, 0),
// on the side of y of the given sign.
inline void place(double dA, double dB, double l, double sign, double& x, double& y)
{
    x = (dA * dA - dB * dB + l * l) / (2.0 * l);
    y = sign * std::sqrt(std::fabs(dA * dA - x * x));
}

}  // namespace

std::string trace_arc(const HostMesh& m, int u, int v, int srcEdge, double srcX, int dstEdge, int hopCap,
                      std::vector<double>& pts, double* dstXTraced)
{
    if (u < 0 || v < 0 || u >= m.nv || v >= m.nv) return "vertex out of range";
    if (srcEdge == -1 && dstEdge == -1) {
        // no crossing: u and v share a mesh edge
        for (int k = 0; k < 3; ++k) pts.push_back(m.xyz[3 * v + k]);
        if (dstXTraced) *dstXTraced = 0.0;
        return std::string();
    }
    if (srcEdge < 0 || dstEdge < 0 || srcEdge >= m.ne || dstEdge >= m.ne) return "key point edge out of range";
    // first face: the one across srcEdge from u's face, with u unfolded below it
    int e = srcEdge;
    const int te = m.twin[e];
    if (te < 0) return "first key point on a boundary edge";
    if (m.tri[3 * (te / 3) + (te % 3 + 2) % 3] != u) return "first key point is not opposite the source";
    double l = m.len[e];
    double sx, sy;
    place(m.len[next_he(te)], m.len[prev_he(te)], l, -1.0, sx, sy);
    double cx = srcX;
    double c3[3];
    for (int hop = 0;; ++hop) {
        key_point_xyz(m, e, cx, c3);
        pts.insert(pts.end(), c3, c3 + 3);
        if (e == dstEdge) break;
        if (hop >= hopCap) return "hop cap reached before the last key point";
        // the far vertex W of e's face, and the ray from the source through c
        double wx, wy;
        place(m.len[prev_he(e)], m.len[next_he(e)], l, 1.0, wx, wy);
        const double dx = cx - sx, dy = -sy;
        const double cross = dx * (wy - sy) - dy * (wx - sx);
        int out;                // the half-edge of this face the ray leaves through
        double px, py;          // exit point
        if (cross > 0.0) {
            // W left of the ray: out through B -> W
            out = next_he(e);
            const double ex = wx - l, ey = wy;
            const double den = dx * ey - dy * ex;
            const double s = ((l - sx) * dy - (0.0 - sy) * dx) / den;
            px = l + s * ex; py = s * ey;
        } else if (cross < 0.0) {
            // out through W -> A
            out = prev_he(e);
            const double ex = -wx, ey = -wy;
            const double den = dx * ey - dy * ex;
            const double s = ((wx - sx) * dy - (wy - sy) * dx) / den;
            px = wx + s * ex; py = wy + s * ey;
        } else {
            return "the geodesic meets a vertex before its end";
        }
        const int ne = m.twin[out];
        if (ne < 0) return "the geodesic leaves the mesh";
        // next frame: entry half-edge ne, whose verts[0] is verts[1](out)
        double n0x, n0y, n1x, n1y;     // its endpoints in this frame
        if (out == next_he(e)) { n0x = wx; n0y = wy; n1x = l; n1y = 0.0; }
        else { n0x = 0.0; n0y = 0.0; n1x = wx; n1y = wy; }
        const double dA = std::hypot(sx - n0x, sy - n0y), dB = std::hypot(sx - n1x, sy - n1y);
        cx = std::hypot(px - n0x, py - n0y);
        e = ne;
        l = m.len[e];
        place(dA, dB, l, -1.0, sx, sy);
    }
    if (dstXTraced) *dstXTraced = cx;
    // the last face's far vertex is v
    if (m.tri[3 * (e / 3) + (e % 3 + 2) % 3] != v) return "last key point is not opposite the destination";
    for (int k = 0; k < 3; ++k) pts.push_back(m.xyz[3 * v + k]);
    return std::string();
}

}  // namespace svg
