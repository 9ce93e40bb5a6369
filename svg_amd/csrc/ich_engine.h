// Single-source bounded ICH window propagation and geodesic backtrace,
// written once over a storage policy P so that the plain (reference-layout)
// and batched (compact SoA) kernels run the same arithmetic in the same
// order.  Every floating-point expression keeps the operand order of the
// reference so that, compiled with -ffp-contract=off, the results are the
// correctly rounded values the CPU restatement produces.
//
// Reference: ICHDevice.cuh:16-93 (records), ICHDevice.cu:67-132 (Execute),
// :336-480 (BuildGeodesicPathTo), :525-581 (Initialize, vertex source),
// :618-726 (PropagateWindow), :728-904 (pseudo-source windows),
// :906-979 (IsValidWindow, BuildWindow, Intersect); Geometry.cu:306-311
// (toLeft).  On open meshes (which the reference cannot build) a boundary
// vertex is never a pseudo-source, and a boundary source also reaches the
// 1-ring vertex across its incoming boundary edge; a 1-ring vertex reached
// over an outgoing boundary half-edge e records e itself as its enter edge,
// and the backtrace takes verts[1](enter) == vertex as a step along e.
//
// The policy P provides
//   mesh:   int twin(e), next(e), prev(e), v0(e), v1(e); double len(e);
//           int first_edge(v); uint8 vflag(v); double cornerA(e), cornerB(e);
//           int nv(), nf();
//           void face3(e0, Tri&) -- the triangle of half-edge e0 as seen from e0
//   tables: VInfo vget(v); void vput(v, VInfo); void vtouch(v); int birth_of(v);
//           SInfo sget(e); void sput(e, SInfo);
//           void lookup_face(Tri, VInfo& a, VInfo& b, VInfo& op, SInfo& split(e0))
//   heaps:  bool wempty(); void wtop(Win&); void wpop(); void wpush(const Win&);
//           bool pempty(); void ptop(Pseudo&); void ppop(); void ppush(const Pseudo&)
//   counters: void count(int which, long long n)
// The stepped form (next_event / do_event) instead pops through
// void pop_either(bool pseudo) and takes the propagation's lookups in two
// halves: a type LookPre, void lookup_issue(Tri, LookPre&) before the pop and
// void lookup_finish(Tri, LookPre&, VInfo& a, VInfo& b, VInfo& op, SInfo&)
// after it; void push_hint() is called before a propagation's pushes.
// A policy with kFanRecords also provides the fan of each vertex by
// position (fan_ok, fan_base, fan_index, fan_cA, fan_cB, fan_window,
// fan_neigh) and a vertex lookup in two halves (VProbe, vprobe_issue,
// vprobe_finish), which gen_sub_wins uses for interior vertices.
#pragma once

#include <hip/hip_runtime.h>
#include <cfloat>
#include <cstdint>

#include "svg_acos.h"

namespace svg {

#define SVG_HD __host__ __device__ __forceinline__

constexpr double kPI = 3.141592653589793238;     // Geometry.cuh:18
constexpr double kRelErr = 1e-8;                 // ICHDevice.cuh:10
constexpr int kHeapCap = 1023;                   // WINPQ_SIZE - 1 (SVG.cuh:13; SVG.cu:19-22)

// Event counters kept per source (units of the algorithmic-bytes model).
enum Counter {
    CNT_WIN_PUSH = 0, CNT_WIN_POP, CNT_PSEUDO_PUSH, CNT_PSEUDO_POP, CNT_PROPAGATE,
    CNT_CHILD_BUILT, CNT_SPLIT_WRITE, CNT_VERT_WRITE, CNT_FAN_HALFEDGE, CNT_PSEUDO_NEIGH,
    CNT_WIN_DROP, CNT_PSEUDO_DROP, CNT_BT_HOPS, CNT_REACHED, CNT_ARCS, CNT_ERRORS,
    CNT_NUM
};

// Window (ICHDevice.cuh:16-50) in registers.  sx, sy cache FlatenedSrc().
struct Win {
    double b0, b1, d0, d1, ps, minDist, sx, sy;
    int edge;
    int twin;                // twin of edge: the half-edge the window crosses into
    int pseudoSrc;
    int birth;
    int level;
};

struct Pseudo {       // ICHDevice.cuh:52-63 (srcId is always the source here)
    double dist;
    int vert;
    int pseudoSrc;
    int birth;
    int level;
};

struct VInfo {        // ICHDevice.cuh:74-86 (srcId / pseudoSrcId are never read)
    double dist;
    int enter;
    int8_t birth;
    uint8_t isSource;
};

struct SInfo {        // ICHDevice.cuh:65-72
    double dist, x;
    int ps;
};

SVG_HD VInfo vinfo_default() { VInfo v; v.dist = DBL_MAX; v.enter = -1; v.birth = -1; v.isSource = 0; return v; }
SVG_HD SInfo sinfo_default() { SInfo s; s.dist = DBL_MAX; s.x = DBL_MAX; s.ps = -1; return s; }

// The triangle on the other side of a crossing: e0 and its two successors.
struct Tri {
    int e0, e1, e2;
    int a, b, op;            // v0(e0), v1(e0), v1(e1) = opposite vertex
    double l0, l1, l2;
    int tw1, tw2;            // twins of e1, e2
};

SVG_HD double len2(double x, double y) { return sqrt(x * x + y * y); }

// Window::calcMinDist + Window::FlatenedSrc (ICHDevice.cuh:30-49)
SVG_HD void win_finish(Win& w)
{
    double wLen = w.b1 - w.b0;
    double xProj = (w.d0 * w.d0 + wLen * wLen - w.d1 * w.d1) / (2.0 * wLen);
    double y = sqrt(fabs(w.d0 * w.d0 - xProj * xProj));
    if (xProj < 0.0) w.minDist = w.d0 + w.ps;
    else if (xProj > wLen) w.minDist = w.d1 + w.ps;
    else w.minDist = y + w.ps;
    w.sx = xProj + w.b0;
    w.sy = y;
}

// ICHDevice::Intersect (ICHDevice.cu:975-979)
SVG_HD double intersect(double v0x, double v0y, double v1x, double v1y,
                        double p0x, double p0y, double p1x, double p1y)
{
    double a00 = p0x - p1x, a01 = v1x - v0x, b0 = v1x - p1x;
    double a10 = p0y - p1y, a11 = v1y - v0y, b1 = v1y - p1y;
    return (b0 * a11 - b1 * a01) / (a00 * a11 - a10 * a01);
}

SVG_HD bool beaten(double a, double b, double tol) { return a > b && a / b - 1.0 > tol; }

// One face crossing of the backtrace unfolding (ICHDevice.cu:420-462).
struct Unfold { int gid; double gpos, lpx, lpy, cpx, cpy; };

template <class P>
struct Ich {
    P& p;
    int source;
    int total;              // totalCalcVertNum
    int K;
    bool pathPassVert;
    int srcFace;            // sourcePointFace: -1 for a vertex source
    bool stalled;           // execute_point stopped on unordered (NaN) queue keys

    SVG_HD Ich(P& pol, int src, int K_)
        : p(pol), source(src), total(0), K(K_), pathPassVert(false), srcFace(-1), stalled(false) {}

    SVG_HD void push_win(const Win& w) { p.count(CNT_WIN_PUSH, 1); p.wpush(w); }
    SVG_HD void push_pseudo(const Pseudo& q) { p.count(CNT_PSEUDO_PUSH, 1); p.ppush(q); }

    // ICHDevice::Initialize, vertex source (ICHDevice.cu:527-581), plus the
    // boundary-source rule above
    SVG_HD void initialize()
    {
        const int src = source;
        const int start = p.first_edge(src);
        // a vertex that no face uses (firstEdge -1, Mesh.cuh:22) starts no
        // window: its row stays empty and no other source reaches it
        if (start == -1) return;
        int cur = start;
        do {
            const int opEdge = p.next(cur);
            Win w;
            w.edge = opEdge;
            w.b0 = 0.0; w.b1 = p.len(opEdge);
            w.d0 = p.len(cur);
            w.d1 = p.len(p.prev(cur));
            w.ps = 0.0;
            win_finish(w);
            w.pseudoSrc = src; w.birth = 0; w.level = 0;
            w.twin = p.twin(opEdge);
            push_win(w);
            const int opVert = p.v1(cur);
            p.vtouch(opVert);
            if (p.len(cur) > p.vget(opVert).dist) {
                cur = p.twin(cur);
                if (cur != -1) cur = p.next(cur);
                continue;
            }
            VInfo vi = p.vget(opVert);
            vi.birth = 0;
            vi.dist = p.len(cur);
            vi.enter = p.twin(cur) != -1 ? p.twin(cur) : cur;     // boundary: e itself
            p.vput(opVert, vi);
            if (p.vflag(opVert) & 1) {                            // angle < 2*PI
                cur = p.twin(cur);
                if (cur != -1) cur = p.next(cur);
                continue;
            }
            Pseudo q;
            q.vert = opVert; q.dist = p.len(cur);
            q.pseudoSrc = src;
            q.birth = p.vget(opVert).birth;
            q.level = 0;
            push_pseudo(q);
            cur = p.twin(cur);
            if (cur != -1) cur = p.next(cur);
        } while (cur != start && cur != -1);
        if (cur == -1) {                                          // boundary source
            const int inE = p.prev(start);
            const int w = p.v0(inE);
            p.vtouch(w);
            if (!(p.len(inE) > p.vget(w).dist)) {
                VInfo vi = p.vget(w);
                vi.birth = 0; vi.dist = p.len(inE); vi.enter = inE;
                p.vput(w, vi);
            }
        }
        VInfo sv = p.vget(src);
        sv.birth = 0; sv.dist = 0.0; sv.enter = -1; sv.isSource = 1;
        p.vput(src, sv);
    }

    // ICHDevice::BuildWindow (ICHDevice.cu:957-973); (sx, sy) is the father's source.
    SVG_HD void build_window(const Win& fa, int edge, double elen, double t0, double t1,
                             double v0x, double v0y, double v1x, double v1y, Win& w)
    {
        w.edge = edge;
        w.b0 = (1 - t0) * elen;
        w.b1 = (1 - t1) * elen;
        {
            double mx = t0 * v0x + (1 - t0) * v1x, my = t0 * v0y + (1 - t0) * v1y;
            w.d0 = len2(fa.sx - mx, fa.sy - my);
        }
        {
            double mx = t1 * v0x + (1 - t1) * v1x, my = t1 * v0y + (1 - t1) * v1y;
            w.d1 = len2(fa.sx - mx, fa.sy - my);
        }
        w.ps = fa.ps;
        win_finish(w);
        w.pseudoSrc = fa.pseudoSrc;
        w.birth = fa.birth;
        w.level = fa.level + 1;
    }

    // ICHDevice::IsValidWindow (ICHDevice.cu:906-955).  The child lies on an
    // edge whose verts[0], verts[1] and opposite vertex have distances dv1,
    // dv2, dv3; l0 = len(edge), l1 = len(next), l2 = len(next next).
    SVG_HD bool is_valid(const Win& w, bool isLeft, double dv1, double dv2, double dv3,
                         double l0, double l1, double l2)
    {
        if (w.b1 <= w.b0) return false;
        double p3x = (l2 * l2 + l0 * l0 - l1 * l1) / (2.0 * l0);
        double p3y = sqrt(fabs(l2 * l2 - p3x * p3x));
        double ax = w.b0, bx = w.b1;
        double srcB = w.ps + len2(w.sx - bx, w.sy - 0.0);
        double srcA = w.ps + len2(w.sx - ax, w.sy - 0.0);
        if (beaten(srcB, dv1 + w.b1, 0.0)) return false;
        if (beaten(srcA, dv2 + l0 - w.b0, 0.0)) return false;
        if (isLeft) {
            if (beaten(srcA, dv3 + len2(p3x - ax, p3y - 0.0), 0.0)) return false;
        } else {
            if (beaten(srcB, dv3 + len2(p3x - bx, p3y - 0.0), kRelErr)) return false;
        }
        return true;
    }

    // ICHDevice::PropagateWindow (ICHDevice.cu:618-726).  The three vertex
    // records of the far triangle and its split record are read together up
    // front; the only one the step can change before IsValidWindow reads it
    // again is the opposite vertex, whose new distance is then `cand`.
    SVG_HD void propagate(const Win& win) { Tri t; propagate(win, false, t); }
    // faceRead: `face` already holds the far face of the window's twin edge.
    SVG_HD void propagate(const Win& win, bool faceRead, const Tri& face)
    {
        const int e0 = win.twin;
        if (e0 == -1) return;
        Tri t;
        if (faceRead) t = face;
        else p.face3(e0, t);
        VInfo va, vb, vo;
        SInfo sp;
        p.push_hint();
        p.lookup_face(t, va, vb, vo, sp);
        propagate_body(win, t, va, vb, vo, sp);
    }
    // the table lookups of the far face were issued before the pop
    template <class Pre>
    SVG_HD void propagate_pre(const Win& win, const Tri& t, Pre& pre)
    {
        VInfo va, vb, vo;
        SInfo sp;
        p.push_hint();
        p.lookup_finish(t, pre, va, vb, vo, sp);
        propagate_body(win, t, va, vb, vo, sp);
    }
    SVG_HD void propagate_body(const Win& win, const Tri& t, const VInfo& va, const VInfo& vb, const VInfo& vo0,
                               const SInfo& sp)
    {
        const int e0 = win.twin;
        p.count(CNT_PROPAGATE, 1);
        VInfo vo = vo0;
        const double l0 = t.l0, l1 = t.l1, l2 = t.l2;
        const double sx = win.sx, sy = win.sy;
        const double leftx = win.b0, rightx = win.b1;
        const double v0x = 0.0, v0y = 0.0, v1x = l0, v1y = 0.0;
        const double v2x = (l1 * l1 + l0 * l0 - l2 * l2) / (2.0 * l0);
        const double v2y = -sqrt(fabs(l1 * l1 - v2x * v2x));
        const double interX = v2x - v2y * (v2x - sx) / (v2y - sy);
        Win lw, rw;
        bool hasL = true, hasR = true;
        double dOp = vo.dist;
        // the ray from the source through the far vertex passes left of the
        // window (only the right child), right of it (only the left child),
        // or through it (both, after the split and vertex updates)
        const bool onlyR = interX <= leftx;
        const bool onlyL = !onlyR && interX >= rightx;
        if (onlyR) {
            hasL = false;
        } else if (onlyL) {
            hasR = false;
        } else {
            const double directDist = len2(v2x - sx, v2y - sy);
            const double cand = directDist + win.ps;
            if (cand > sp.dist && cand / sp.dist - 1.0 > kRelErr) {
                hasL = sp.x < interX;
                hasR = !hasL;
            } else {
                if (cand < sp.dist) {
                    SInfo ns;
                    ns.dist = cand;
                    ns.ps = win.pseudoSrc;
                    ns.x = l0 - interX;
                    p.sput(e0, ns);
                    p.count(CNT_SPLIT_WRITE, 1);
                }
                VInfo ov = vo;
                if (cand < ov.dist) {
                    if (ov.dist == DBL_MAX) ++total;
                    ov.birth = (int8_t)(ov.birth + 1);
                    ov.dist = cand;
                    ov.enter = e0;
                    p.vput(t.op, ov);
                    dOp = cand;
                    p.count(CNT_VERT_WRITE, 1);
                    if (p.vflag(t.op) & 2) {                        // angle > 2*PI
                        Pseudo q;
                        q.vert = t.op; q.dist = ov.dist;
                        q.pseudoSrc = win.pseudoSrc;
                        q.birth = ov.birth;
                        q.level = win.level + 1;
                        push_pseudo(q);
                    }
                }
            }
        }
        // Each child is built by one code path for all three cases (the
        // cases differ only in the clipping parameters), so lanes of a wave
        // in different cases build their children together.
        // child on e1 (left): edge e1 runs b -> op; its next is e2 (op -> a)
        if (hasL) {
            const double t0 = intersect(sx, sy, leftx, 0.0, v0x, v0y, v2x, v2y);
            const double t1 = onlyL ? intersect(sx, sy, rightx, 0.0, v0x, v0y, v2x, v2y) : 0.0;
            build_window(win, t.e1, l1, t0, t1, v0x, v0y, v2x, v2y, lw);
            p.count(CNT_CHILD_BUILT, 1);
            if (!is_valid(lw, true, vb.dist, dOp, va.dist, l1, l2, l0)) hasL = false;
        }
        // child on e2 (right): edge e2 runs op -> a; its next is e0 (a -> b)
        if (hasR) {
            const double t0 = onlyR ? intersect(sx, sy, leftx, 0.0, v2x, v2y, v1x, v1y) : 1.0;
            const double t1 = intersect(sx, sy, rightx, 0.0, v2x, v2y, v1x, v1y);
            build_window(win, t.e2, l2, t0, t1, v2x, v2y, v1x, v1y, rw);
            p.count(CNT_CHILD_BUILT, 1);
            if (!is_valid(rw, false, dOp, va.dist, vb.dist, l2, l0, l1)) hasR = false;
        }
        if (hasL) { lw.twin = t.tw1; push_win(lw); }
        if (hasR) { rw.twin = t.tw2; push_win(rw); }
    }

    // The two starting angles of GenSubWinsForPseudoSrcFromWindow
    // (ICHDevice.cu:795-825): the pseudo-source is verts[1](next(e0)), the
    // window's split point lies on e0.
    SVG_HD void window_sweep_angles(int e0, double& angle0, double& angle1)
    {
        const int e1 = p.next(e0);
        const int e2 = p.next(e1);
        const double l0 = p.len(e0), l1 = p.len(e1), l2 = p.len(e2);
        const double epx = l0 - p.sget(e0).x, epy = 0.0;
        const double v0x = 0.0, v0y = 0.0, v1x = l0, v1y = 0.0;
        const double v2x = (l1 * l1 + l0 * l0 - l2 * l2) / (2.0 * l0);
        const double v2y = -sqrt(fabs(l1 * l1 - v2x * v2x));
        const double dx = epx - v2x, dy = epy - v2y;
        const double dlen = len2(dx, dy);
        double a0 = (dx * (v0x - v2x) + dy * (v0y - v2y)) / dlen / l1;
        double a1 = (dx * (v1x - v2x) + dy * (v1y - v2y)) / dlen / l2;
        if (a0 > 1.0) a0 = 1.0; else if (a0 < -1.0) a0 = -1.0;
        if (a1 > 1.0) a1 = 1.0; else if (a1 < -1.0) a1 = -1.0;
        angle0 = svg_acos(a0); angle1 = svg_acos(a1);
    }

    // ICHDevice::GenSubWinsForPseudoSrcFromWindow (ICHDevice.cu:795-858)
    SVG_HD void sweep_from_window(const Pseudo& pw, int e0, int& startEdge, int& endEdge)
    {
        (void)pw;
        const int e1 = p.next(e0);
        const int e2 = p.next(e1);
        double angle0, angle1;
        window_sweep_angles(e0, angle0, angle1);
        startEdge = -1; endEdge = -1;
        int cur = p.twin(e1);
        while (angle0 < kPI && cur != -1) {
            angle0 += p.cornerA(cur);
            p.count(CNT_FAN_HALFEDGE, 1);
            cur = p.twin(p.prev(cur));
        }
        if (cur != -1) startEdge = p.next(p.twin(cur));
        cur = p.twin(e2);
        while (angle1 < kPI && cur != -1) {
            angle1 += p.cornerB(cur);
            p.count(CNT_FAN_HALFEDGE, 1);
            cur = p.twin(p.next(cur));
        }
        if (cur != -1) endEdge = p.twin(p.prev(p.twin(cur)));
    }

    // ICHDevice::GenSubWinsForPseudoSrcFromPseudoSrc (ICHDevice.cu:860-904)
    SVG_HD void sweep_from_pseudo(int enter, int& startEdge, int& endEdge)
    {
        double angle0 = 0.0, angle1 = 0.0;
        startEdge = -1; endEdge = -1;
        int cur = enter;
        while (angle0 < kPI && cur != -1) {
            angle0 += p.cornerA(cur);
            p.count(CNT_FAN_HALFEDGE, 1);
            cur = p.twin(p.prev(cur));
        }
        if (cur != -1) startEdge = p.next(p.twin(cur));
        cur = p.twin(enter);
        while (angle1 < kPI && cur != -1) {
            angle1 += p.cornerB(cur);
            p.count(CNT_FAN_HALFEDGE, 1);
            cur = p.twin(p.next(cur));
        }
        if (cur != -1) endEdge = p.twin(p.prev(p.twin(cur)));
    }

    // ICHDevice::GenSubWinsForPseudoSrc (ICHDevice.cu:728-904) over the fan
    // records of a policy that keeps them (kFanRecords): the same sweeps,
    // windows and neighbour updates, in the same order and arithmetic, with
    // the fan addressed by position instead of walked one dependent read at
    // a time.  Position j - 1 (cyclically) is twin(prev(o_j)), the step of
    // the first sweep and of the window loop; the second sweep's incoming
    // half-edges twin(o_j) advance with j + 1.  Only interior vertices are
    // pseudo-sources, so every fan here is a full cycle.
    SVG_HD void gen_sub_wins_fan(const Pseudo& pw)
    {
        const int vert = pw.vert;
        const VInfo vi = p.vget(vert);
        int base, d;
        p.fan_base(vert, base, d);
        const int last = base + d - 1;
        int jS = base, jE = base;         // first window position, stop position
        const int enter = vi.enter;
        if (!(enter == -1 && vi.birth != -1)) {
            int jA, jB;                   // where the two sweeps start
            double angle0 = 0.0, angle1 = 0.0;
            if (enter >= 0 && p.v0(enter) == vert) {                    // from a pseudo-source
                jA = p.fan_index(enter);
                jB = jA;
            } else if (enter >= 0 && p.v1(p.next(enter)) == vert) {     // from a window
                window_sweep_angles(enter, angle0, angle1);
                jB = p.fan_index(p.next(p.next(enter)));
                jA = jB == base ? last : jB - 1;
            } else {
                p.count(CNT_ERRORS, 1);       // unreachable: only interior saddles are pseudo-sources
                return;
            }
            // first sweep: cornerA at jA, jA - 1, ... until the angle reaches
            // pi; four positions are read at once (those past the stop unused)
            int j = jA;
            for (;;) {
                const int j1 = j == base ? last : j - 1;
                const int j2 = j1 == base ? last : j1 - 1;
                const int j3 = j2 == base ? last : j2 - 1;
                const double c0 = p.fan_cA(j), c1 = p.fan_cA(j1), c2 = p.fan_cA(j2), c3 = p.fan_cA(j3);
                if (!(angle0 < kPI)) break;
                angle0 += c0; p.count(CNT_FAN_HALFEDGE, 1); j = j1;
                if (!(angle0 < kPI)) break;
                angle0 += c1; p.count(CNT_FAN_HALFEDGE, 1); j = j2;
                if (!(angle0 < kPI)) break;
                angle0 += c2; p.count(CNT_FAN_HALFEDGE, 1); j = j3;
                if (!(angle0 < kPI)) break;
                angle0 += c3; p.count(CNT_FAN_HALFEDGE, 1); j = j3 == base ? last : j3 - 1;
            }
            jS = j == last ? base : j + 1;                 // next(twin(o_j)) = o_{j+1}
            // second sweep: cornerB of twin(o_j) at jB, jB + 1, ...
            j = jB;
            for (;;) {
                const int j1 = j == last ? base : j + 1;
                const int j2 = j1 == last ? base : j1 + 1;
                const int j3 = j2 == last ? base : j2 + 1;
                const double c0 = p.fan_cB(j), c1 = p.fan_cB(j1), c2 = p.fan_cB(j2), c3 = p.fan_cB(j3);
                if (!(angle1 < kPI)) break;
                angle1 += c0; p.count(CNT_FAN_HALFEDGE, 1); j = j1;
                if (!(angle1 < kPI)) break;
                angle1 += c1; p.count(CNT_FAN_HALFEDGE, 1); j = j2;
                if (!(angle1 < kPI)) break;
                angle1 += c2; p.count(CNT_FAN_HALFEDGE, 1); j = j3;
                if (!(angle1 < kPI)) break;
                angle1 += c3; p.count(CNT_FAN_HALFEDGE, 1); j = j3 == last ? base : j3 + 1;
            }
            jE = j == base ? last : j - 1;                 // twin(prev(twin(twin(o_j)))) = o_{j-1}
        }
        // windows over the edges opposite the fan positions jS, jS - 1, ...
        // down to (not including) jE; the next record is read before a push,
        // and the heap ancestors of the first two push slots (and the pseudo
        // heap's next parent) are read ahead as for a propagation
        {
            p.push_hint();
            int j = jS, guard = 0;
            double lo, lnx, lpv;
            int o, twn;
            p.fan_window(j, lo, lnx, lpv, o, twn);
            for (;;) {
                const int jn = j == base ? last : j - 1;
                double lo2, lnx2, lpv2;
                int o2, twn2;
                p.fan_window(jn, lo2, lnx2, lpv2, o2, twn2);
                Win w;
                w.edge = p.next(o);
                w.b0 = 0.0; w.b1 = lnx;
                w.d0 = lo;
                w.d1 = lpv;
                w.ps = pw.dist;
                win_finish(w);
                w.pseudoSrc = vert;
                w.birth = pw.birth;
                w.level = pw.level + 1;
                w.twin = twn;
                push_win(w);
                j = jn;
                if (++guard > 4096) { p.count(CNT_ERRORS, 1); break; }
                if (j == jE) break;
                lo = lo2; lnx = lnx2; lpv = lpv2; o = o2; twn = twn2;
            }
        }
        // neighbours in fan order from firstEdge, their first probes issued
        // two at a time (the neighbours are distinct vertices, and a probe
        // that meets another key continues on the current table)
        for (int k0 = 0; k0 < d; k0 += 2) {
            int nb[2], tw[2], fl[2];
            double ln[2];
            typename P::VProbe q[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) p.fan_neigh(base + (k0 + i < d ? k0 + i : d - 1), nb[i], tw[i], fl[i], ln[i]);
#pragma unroll
            for (int i = 0; i < 2; ++i) p.vprobe_issue(nb[i], q[i]);
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                if (k0 + i < d) {
                    p.count(CNT_PSEUDO_NEIGH, 1);
                    VInfo ov = p.vprobe_finish(nb[i], q[i]);
                    if (!((fl[i] & 1) || ov.dist < pw.dist + ln[i])) {
                        if (ov.dist == DBL_MAX) ++total;
                        ov.dist = pw.dist + ln[i];
                        ov.birth = (int8_t)(ov.birth + 1);
                        ov.enter = tw[i];
                        p.vput(nb[i], ov);
                        p.count(CNT_VERT_WRITE, 1);
                        Pseudo c;
                        c.vert = nb[i]; c.dist = ov.dist;
                        c.pseudoSrc = vert;
                        c.birth = ov.birth;
                        c.level = pw.level;
                        push_pseudo(c);
                    }
                }
            }
        }
    }

    // ICHDevice::GenSubWinsForPseudoSrc (ICHDevice.cu:728-793)
    SVG_HD void gen_sub_wins(const Pseudo& pw)
    {
        if constexpr (P::kFanRecords) {
            if (p.fan_ok() && !(p.vflag(pw.vert) & 4)) { gen_sub_wins_fan(pw); return; }
        }
        int startEdge = -1, endEdge = -1;
        const VInfo vi = p.vget(pw.vert);
        if (vi.enter == -1 && vi.birth != -1) {
            startEdge = p.first_edge(pw.vert);
            endEdge = startEdge;
        } else if (vi.enter >= 0 && p.v0(vi.enter) == pw.vert) {
            sweep_from_pseudo(vi.enter, startEdge, endEdge);
        } else if (vi.enter >= 0 && p.v1(p.next(vi.enter)) == pw.vert) {
            sweep_from_window(pw, vi.enter, startEdge, endEdge);
        } else {
            p.count(CNT_ERRORS, 1);      // unreachable: only interior saddles are pseudo-sources
            return;
        }
        if (startEdge == -1 || endEdge == -1) { p.count(CNT_ERRORS, 1); return; }
        int guard = 0;
        do {
            Win w;
            w.edge = p.next(startEdge);
            w.b0 = 0.0; w.b1 = p.len(w.edge);
            w.d0 = p.len(startEdge);
            w.d1 = p.len(p.next(w.edge));
            w.ps = pw.dist;
            win_finish(w);
            w.pseudoSrc = pw.vert;
            w.birth = pw.birth;
            w.level = pw.level + 1;
            w.twin = p.twin(w.edge);
            push_win(w);
            startEdge = p.twin(p.prev(startEdge));
            if (startEdge == -1 || ++guard > 4096) { p.count(CNT_ERRORS, 1); break; }
        } while (startEdge != endEdge);

        // The mesh data of the next fan edge is read while this neighbour is
        // looked up (the mesh never changes, so the order of the table
        // reads and writes is the reference's).
        const int start = p.first_edge(pw.vert);
        int cur = start;
        int opVert = p.v1(cur);
        double len = p.len(cur);
        int tw = p.twin(cur);
        do {
            const int nxt = tw != -1 ? p.next(tw) : -1;
            int opVert2 = 0, tw2 = -1;
            double len2 = 0.0;
            if (nxt != -1 && nxt != start) { opVert2 = p.v1(nxt); len2 = p.len(nxt); tw2 = p.twin(nxt); }
            p.count(CNT_PSEUDO_NEIGH, 1);
            VInfo ov = p.vget(opVert);
            if (!((p.vflag(opVert) & 1) || ov.dist < pw.dist + len)) {
                if (ov.dist == DBL_MAX) ++total;
                ov.dist = pw.dist + len;
                ov.birth = (int8_t)(ov.birth + 1);
                ov.enter = tw;
                p.vput(opVert, ov);
                p.count(CNT_VERT_WRITE, 1);
                Pseudo c;
                c.vert = opVert; c.dist = ov.dist;
                c.pseudoSrc = pw.vert;
                c.birth = ov.birth;
                c.level = pw.level;
                push_pseudo(c);
            }
            cur = nxt; opVert = opVert2; len = len2; tw = tw2;
        } while (cur != start && cur != -1);
    }

    // ICHDevice::Execute (ICHDevice.cu:67-132) for a vertex source.
    SVG_HD void execute()
    {
        initialize();
        const int nf = p.nf();
        while (!p.wempty() || !p.pempty()) {
            Win top;
            bool haveTop = false;
            while (!p.wempty()) {
                p.wtop(top);
                if (top.birth == p.birth_of(top.pseudoSrc)) { haveTop = true; break; }
                p.wpop();
                p.count(CNT_WIN_POP, 1);
            }
            Pseudo ptop;
            bool havePtop = false;
            while (!p.pempty()) {
                p.ptop(ptop);
                if (ptop.birth == p.birth_of(ptop.vert)) { havePtop = true; break; }
                p.ppop();
                p.count(CNT_PSEUDO_POP, 1);
            }
            if (haveTop && (!havePtop || top.minDist < ptop.dist)) {
                p.wpop();
                p.count(CNT_WIN_POP, 1);
                if (top.level > nf) continue;
                propagate(top);
            } else if (havePtop && (!haveTop || top.minDist >= ptop.dist)) {
                p.ppop();
                p.count(CNT_PSEUDO_POP, 1);
                if (ptop.level >= nf) continue;
                gen_sub_wins(ptop);
            } else {
                // the tops compare unordered (a NaN key): the reference would
                // spin on them (ICHDevice.cu:88, 110); the source ends flagged
                p.count(CNT_ERRORS, 1);
                break;
            }
            if (K != -1 && total >= K) break;
            if (p.stop()) break;
        }
    }

    // Stepped form of execute() for lanes that share a wave.  next_event()
    // names the next heap operation of this source without changing state
    // and do_event() performs it, so a source runs exactly the operations of
    // execute(), in its order (stale pops are interleaved between the two
    // heaps, which are independent and do not change births), while the
    // caller chooses which lanes of a wave step together.
    enum : int { EV_DONE = 0, EV_STALE_W = 1, EV_STALE_P = 2, EV_WIN = 3, EV_PSEUDO = 4, EV_STUCK = 5 };
    SVG_HD void reset(int src)
    {
        source = src;
        total = 0;
        pathPassVert = false;
    }
    SVG_HD int next_event(Win& top, Pseudo& ptop)
    {
        bool haveTop = false, havePtop = false;
        if (!p.wempty()) {
            p.wtop(top);
            if (top.birth != p.birth_of(top.pseudoSrc)) return EV_STALE_W;
            haveTop = true;
        }
        if (!p.pempty()) {
            p.ptop(ptop);
            if (ptop.birth != p.birth_of(ptop.vert)) return EV_STALE_P;
            havePtop = true;
        }
        if (haveTop && (!havePtop || top.minDist < ptop.dist)) return EV_WIN;
        if (havePtop && (!haveTop || top.minDist >= ptop.dist)) return EV_PSEUDO;
        // unordered keys (NaN): execute() would spin on the same tops
        return haveTop || havePtop ? EV_STUCK : EV_DONE;
    }
    // Performs `ev`; false once the source is finished (K reached or stopped).
    // The policy pops either queue through one call (pop_either: the queues
    // share a layout), so lanes of a wave popping a window, a pseudo-window
    // or a stale entry of either run the pop together.  The far face of a
    // window that will propagate is read, and its table lookups issued
    // (lookup_issue), before the pop, so those reads overlap the sift-down.
    SVG_HD bool do_event(int ev, const Win& top, const Pseudo& ptop)
    {
        const int nf = p.nf();
        if (ev < EV_STALE_W || ev > EV_PSEUDO) return false;
        const bool isP = ev == EV_STALE_P || ev == EV_PSEUDO;
        const bool prop = ev == EV_WIN && !(top.level > nf) && top.twin != -1;
        Tri t;
        typename P::LookPre pre;
        if (prop) { p.face3(top.twin, t); p.lookup_issue(t, pre); }
        p.pop_either(isP);
        if (isP) p.count(CNT_PSEUDO_POP, 1);
        else p.count(CNT_WIN_POP, 1);
        if (ev == EV_WIN) {
            if (!prop) return true;
            propagate_pre(top, t, pre);
        } else if (ev == EV_PSEUDO) {
            if (ptop.level >= nf) return true;
            gen_sub_wins(ptop);
        } else {
            return true;
        }
        if (K != -1 && total >= K) return false;
        return !p.stop();
    }

    // One unfolding step (ICHDevice.cu:420-462).
    SVG_HD bool unfold(Unfold& u)
    {
        const int e0 = p.twin(u.gid);
        if (e0 == -1) return false;
        Tri t;
        p.face3(e0, t);
        const double l0 = t.l0, l1 = t.l1, l2 = t.l2;
        const double ox = (l0 * l0 + l2 * l2 - l1 * l1) / (2.0 * l0);
        const double oy = sqrt(fabs(l2 * l2 - ox * ox));
        // toLeft(opVert2D, lastPoint, curPoint) (Geometry.cu:306-311)
        double ax = u.cpx - u.lpx, ay = u.cpy - u.lpy;
        double bx = ox - u.lpx, by = oy - u.lpy;
        const double la = len2(ax, ay);
        if (!(fabs(la) < 1e-12)) { ax /= la; ay /= la; }
        const double lb = len2(bx, by);
        if (!(fabs(lb) < 1e-12)) { bx /= lb; by /= lb; }
        const bool left = (ax * by - ay * bx) > 0.0;
        const double f = u.gpos / l0;
        if (left) {
            const double p0x = (l2 * l2 + l1 * l1 - l0 * l0) / (2.0 * l1);
            const double p0y = -sqrt(fabs(l2 * l2 - p0x * p0x));
            const double p1x = l1, p1y = 0.0;
            const double nlx = f * p0x + (1.0 - f) * p1x, nly = f * p0y + (1.0 - f) * p1y;
            const double tt = intersect(u.lpx, u.lpy, u.cpx, u.cpy, l0, 0.0, ox, oy);
            u.gpos = (1.0 - tt) * l1;
            u.gid = t.e1;
            u.cpx = l1 - u.gpos; u.cpy = 0.0;
            u.lpx = nlx; u.lpy = nly;
        } else {
            const double p0x = 0.0, p0y = 0.0;
            const double p1x = (l2 * l2 + l0 * l0 - l1 * l1) / (2.0 * l2);
            const double p1y = -sqrt(fabs(l0 * l0 - p1x * p1x));
            const double nlx = f * p0x + (1.0 - f) * p1x, nly = f * p0y + (1.0 - f) * p1y;
            const double tt = intersect(u.lpx, u.lpy, u.cpx, u.cpy, ox, oy, 0.0, 0.0);
            u.gpos = (1.0 - tt) * l2;
            u.gid = t.e2;
            u.cpx = l2 - u.gpos; u.cpy = 0.0;
            u.lpx = nlx; u.lpy = nly;
        }
        return true;
    }

    // ICHDevice::BuildGeodesicPathTo(vertId, ...) (ICHDevice.cu:336-480), plus
    // steps along a boundary enter edge.
    // Returns pathPassVert; fills the two key points (edge -1 -> x = 0).
    SVG_HD bool path_to(int vertId, int& srcEdge, double& srcX, int& dstEdge, double& dstX, int hopCap)
    {
        bool pass = false;
        srcEdge = -1; dstEdge = -1; srcX = 0.0; dstX = 0.0;
        int curVert = vertId;
        int hops = 0;
        VInfo cv = p.vget(curVert);
        while (cv.dist != 0.0) {
            if (++hops > hopCap) { p.count(CNT_ERRORS, 1); return true; }
            p.count(CNT_BT_HOPS, 1);
            const int enter = cv.enter;
            if (enter == -1) {
                if (dstEdge == -1) { dstEdge = p.first_edge(curVert); dstX = 0.0; }
                else { srcEdge = p.first_edge(curVert); srcX = 0.0; }
                return true;
            } else if (p.v0(enter) == curVert) {
                const int nextVert = p.v1(enter);
                if (!p.vget(nextVert).isSource) {
                    if (dstEdge == -1) { dstEdge = p.next(enter); dstX = 0.0; }
                    else { srcEdge = p.next(enter); srcX = 0.0; }
                    pass = true;
                }
                curVert = nextVert;
            } else if (p.v1(enter) == curVert) {                       // along a boundary enter edge
                const int nextVert = p.v0(enter);
                if (!p.vget(nextVert).isSource) {
                    if (dstEdge == -1) { dstEdge = enter; dstX = 0.0; }
                    else { srcEdge = enter; srcX = 0.0; }
                    pass = true;
                }
                curVert = nextVert;
            } else {
                const SInfo sp = p.sget(enter);
                Unfold u;
                u.gid = enter;
                u.gpos = sp.x;
                if (dstEdge == -1) { dstEdge = u.gid; dstX = u.gpos; }
                else { srcEdge = u.gid; srcX = u.gpos; }
                int tw = p.twin(u.gid);
                if (tw == -1) { p.count(CNT_ERRORS, 1); return true; }
                int opVert = p.v1(p.next(tw));
                const double l0 = p.len(u.gid);
                const double l1 = p.len(p.next(u.gid));
                const double l2 = p.len(p.prev(u.gid));
                u.lpx = (l1 * l1 + l0 * l0 - l2 * l2) / (2.0 * l0);
                u.lpy = -sqrt(fabs(l1 * l1 - u.lpx * u.lpx));
                u.cpx = l0 - u.gpos; u.cpy = 0.0;
                const int ps = sp.ps;
                // a surface-point source's windows carry pseudo-source id V
                const bool pointPs = ps >= p.nv() && srcFace >= 0;
                if (ps < 0 || (ps >= p.nv() && !pointPs)) { p.count(CNT_ERRORS, 1); return true; }
                // unfold to the pseudo-source vertex, or (point source,
                // ICHDevice.cu:415-418) until the face behind is the source's
                while (pointPs ? tw / 3 != srcFace : opVert != ps) {
                    if (++hops > hopCap) { p.count(CNT_ERRORS, 1); return true; }
                    p.count(CNT_BT_HOPS, 1);
                    if (!unfold(u)) { p.count(CNT_ERRORS, 1); return true; }
                    srcEdge = u.gid; srcX = u.gpos;
                    tw = p.twin(u.gid);
                    if (tw == -1) { p.count(CNT_ERRORS, 1); return true; }
                    opVert = p.v1(p.next(tw));
                }
                if (pointPs) return pass;                              // ICHDevice.cu:465-468
                if (!p.vget(opVert).isSource) {
                    srcEdge = p.first_edge(opVert); srcX = 0.0;
                    pass = true;
                }
                curVert = opVert;
            }
            cv = p.vget(curVert);
        }
        return pass;
    }

    // ---------------------------------------------------------------------
    // Surface-point sources.  Half-edge 3f+j is edge j of face f (Mesh.cu:
    // 69-77), so face(e) = e / 3.
    // ---------------------------------------------------------------------

    // ICHDevice::Initialize, surface-point branch (ICHDevice.cu:583-615):
    // a window over each edge of face f with pseudo-source id V (never stale)
    // and the face's vertices at their straight-line distances d[j] (from the
    // point to verts[0] of half-edge 3f+j).
    SVG_HD void initialize_point(int f, const double d[3])
    {
        const int nv = p.nv();
        for (int j = 0; j < 3; ++j) {
            const int opEdge = 3 * f + j;
            Win w;
            w.edge = opEdge;
            w.b0 = 0.0; w.b1 = p.len(opEdge);
            w.d0 = d[j];
            w.d1 = d[j == 2 ? 0 : j + 1];
            w.ps = 0.0;
            win_finish(w);
            w.pseudoSrc = nv; w.birth = 0; w.level = 0;
            w.twin = p.twin(opEdge);
            push_win(w);
            const int opVert = p.v0(opEdge);
            p.vtouch(opVert);
            if (d[j] > p.vget(opVert).dist) continue;
            VInfo vi = p.vget(opVert);
            vi.birth = 0; vi.dist = d[j]; vi.enter = -1;
            p.vput(opVert, vi);
            if (p.vflag(opVert) & 1) continue;                    // angle < 2*PI
            Pseudo q;
            q.vert = opVert; q.dist = d[j];
            q.pseudoSrc = nv;
            q.birth = p.vget(opVert).birth;
            q.level = 0;
            push_pseudo(q);
        }
    }

    // ICHDevice::Execute (ICHDevice.cu:67-132) for a surface-point source.
    // Windows that cross into keptFace are stored as they are popped
    // (ICHDevice.cu:93-107), at most storeCap of them.  The pseudo-source
    // queue's stale test is guarded by the window queue's top, slot 1 of its
    // array, which keeps the last window popped once the queue is empty
    // (ICHDevice.cu:84): while that window descends from the point, stale
    // pseudo-sources are expanded rather than skipped.  guardPs tracks that
    // slot: every pop is of a top read just before.
    SVG_HD void execute_point(Win* stored, int storeCap, int& nStored, int keptFace)
    {
        const int nf = p.nf(), nv = p.nv();
        int guardPs = nv;
        while (!p.wempty() || !p.pempty()) {
            Win top;
            bool haveTop = false;
            while (!p.wempty()) {
                p.wtop(top);
                guardPs = top.pseudoSrc;
                if (top.pseudoSrc >= nv || top.birth == p.birth_of(top.pseudoSrc)) { haveTop = true; break; }
                p.wpop();
                p.count(CNT_WIN_POP, 1);
            }
            Pseudo ptop;
            bool havePtop = false;
            while (!p.pempty()) {
                p.ptop(ptop);
                if (guardPs >= nv || ptop.birth == p.birth_of(ptop.vert)) { havePtop = true; break; }
                p.ppop();
                p.count(CNT_PSEUDO_POP, 1);
            }
            if (haveTop && (!havePtop || top.minDist < ptop.dist)) {
                p.wpop();
                p.count(CNT_WIN_POP, 1);
                if (top.level > nf) continue;
                if (top.minDist != top.minDist) {      // NaN key: stop, as below
                    p.count(CNT_ERRORS, 1);
                    stalled = true;
                    break;
                }
                if (top.twin != -1 && keptFace >= 0 && top.twin / 3 == keptFace && nStored < storeCap)
                    stored[nStored++] = top;
                propagate(top);
            } else if (havePtop && (!haveTop || top.minDist >= ptop.dist)) {
                p.ppop();
                p.count(CNT_PSEUDO_POP, 1);
                if (ptop.level >= nf) continue;
                gen_sub_wins(ptop);
            } else if (haveTop || havePtop) {
                // unordered keys (NaN: a point on a vertex of its face puts
                // the source on a window's line); the reference loops here.
                // The run stops at the first NaN key it meets.
                p.count(CNT_ERRORS, 1);
                stalled = true;
                break;
            }
            if (K != -1 && total >= K) break;
            if (p.stop()) break;
        }
    }

    // ICHDevice::BuildGeodesicPathTo(faceId, pos, ...) (ICHDevice.cu:143-334):
    // distance from the source to a point of face f through a stored window
    // or one of f's vertices, and the key points of that path.  fv[j] is
    // verts[0] of half-edge 3f+j and dP[j] the point's straight-line distance
    // to it.  srcX / dstX keep the caller's values unless a key point is
    // written.  Returns DBL_MAX (no key points) when nothing reaches f.
    SVG_HD double face_path_to(int f, const int fv[3], const double dP[3], const Win* stored, int nStored,
                               int& srcEdge, double& srcX, int& dstEdge, double& dstX, int hopCap)
    {
        const int nv = p.nv();
        srcEdge = -1; dstEdge = -1;
        double minDist = DBL_MAX, xInter = 0.0, posx = 0.0, posy = 0.0;
        int mi = -1, dstVert = -1;
        bool throughWin = true;
        for (int i = 0; i < nStored; ++i) {
            const Win& w = stored[i];
            const int tw = w.twin;
            if (tw == -1 || tw / 3 != f) continue;
            Tri t;
            p.face3(tw, t);
            const double l0 = t.l0, l1 = t.l1, l2 = t.l2;
            const double p0x = 0.0, p0y = 0.0, p1x = l0, p1y = 0.0;
            const double p2x = (l1 * l1 + l0 * l0 - l2 * l2) / (2.0 * l0);
            const double p2y = -sqrt(fabs(l1 * l1 - p2x * p2x));
            // v0 = verts[1](e0), v1 = verts[0](e0), v2 = verts[1](e1)
            double a = 0.0, b = 0.0, c = 0.0;
            for (int j = 0; j < 3; ++j) {
                if (fv[j] == t.b) a = dP[j];
                if (fv[j] == t.a) b = dP[j];
                if (fv[j] == t.op) c = dP[j];
            }
            // the point's barycentric weights from Heron's formula
            double s0 = (b + c + l2) / 2.0;
            double s1 = (a + c + l1) / 2.0;
            double s2 = (a + b + l0) / 2.0;
            s0 = sqrt(fabs(s0 * (s0 - b) * (s0 - c) * (s0 - l2)));
            s1 = sqrt(fabs(s1 * (s1 - a) * (s1 - c) * (s1 - l1)));
            s2 = sqrt(fabs(s2 * (s2 - a) * (s2 - b) * (s2 - l0)));
            const double w0 = s0 / (s0 + s1 + s2);
            const double w1 = s1 / (s0 + s1 + s2);
            const double w2 = s2 / (s0 + s1 + s2);
            const double cx = w0 * p0x + w1 * p1x + w2 * p2x;
            const double cy = w0 * p0y + w1 * p1y + w2 * p2y;
            const double xi = w.sx - (cx - w.sx) / (cy - w.sy) * w.sy;
            if (xi <= 0.0 || xi >= l0) continue;
            double cur;
            if (xi > w.b0 && xi < w.b1) cur = len2(cx - w.sx, cy - w.sy) + w.ps;
            else if (xi <= w.b0) cur = len2(cx - w.b0, cy - 0.0) + w.d0 + w.ps;
            else cur = len2(cx - w.b1, cy - 0.0) + w.d1 + w.ps;
            if (cur < minDist) { minDist = cur; mi = i; xInter = xi; posx = cx; posy = cy; }
        }
        for (int j = 0; j < 3; ++j) {
            const double cd = dP[j] + p.vget(fv[j]).dist;
            if (cd < minDist) { throughWin = false; dstVert = fv[j]; minDist = cd; }
        }
        if (minDist == DBL_MAX) return minDist;
        if (!throughWin) {
            path_to(dstVert, srcEdge, srcX, dstEdge, dstX, hopCap);
            pathPassVert = true;
            dstEdge = p.first_edge(dstVert); dstX = 0.0;
            return minDist;
        }
        const Win& mw = stored[mi];
        Unfold u;
        u.gid = mw.twin;
        u.gpos = p.len(u.gid) - xInter;
        dstEdge = u.gid; dstX = u.gpos;
        int tw = p.twin(u.gid);
        if (tw == -1) { p.count(CNT_ERRORS, 1); return minDist; }
        int opVert = p.v1(p.next(tw));
        u.lpx = posx; u.lpy = posy;
        u.cpx = p.len(u.gid) - u.gpos; u.cpy = 0.0;
        const int ps = mw.pseudoSrc;
        int hops = 0;
        while (ps >= nv ? tw / 3 != srcFace : opVert != ps) {
            if (++hops > hopCap) { p.count(CNT_ERRORS, 1); return minDist; }
            p.count(CNT_BT_HOPS, 1);
            if (!unfold(u)) { p.count(CNT_ERRORS, 1); return minDist; }
            srcEdge = u.gid; srcX = u.gpos;
            tw = p.twin(u.gid);
            if (tw == -1) { p.count(CNT_ERRORS, 1); return minDist; }
            opVert = p.v1(p.next(tw));
        }
        if (ps < nv && !p.vget(opVert).isSource) {
            // the path bends at the pseudo-source: the source-side key point
            // is the one next to the source on that vertex's own path
            pathPassVert = true;
            int de; double dx;
            path_to(opVert, srcEdge, srcX, de, dx, hopCap);
        }
        return minDist;
    }
};

}  // namespace svg
