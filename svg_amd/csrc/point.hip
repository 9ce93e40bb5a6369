// Geodesic distance between two surface points on gfx950 (SVG::SolveSSSD(f0,
// p0, f1, p1), SVG.cu:253-352).
//
// k_point_ich runs the two bounded ICH propagations the query needs, one per
// wave of a 128-thread block (wave 0: from p0, keeping the windows that cross
// into f1; wave 1: from p1), in the reference's table layout
// (plain_policy.h), then wave 0 answers directly through f1
// (BuildGeodesicPathTo(face, pos), ICHDevice.cu:143-334).  The reference runs
// the p0 propagation twice more (SVG.cu:294-296, 330-332) only to rebuild
// state these resident tables still hold.  When the direct answer has no
// source-side key point (SVG.cu:284) the graph stage runs on the solver (a
// multi-source Dijkstra seeded with the p0 run's vertex distances) and
// k_point_join takes the minimum over the p1 run's vertices of graph distance
// plus ICH distance (first table slot on ties, SVG.cu:299-308) and backtraces
// both ends (SVG.cu:309-351).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <cstdio>
#include <cstring>

#include "ich_engine.h"
#include "plain_policy.h"
#include "point.h"

namespace svg {

#define PT_TRY(expr)                                                           \
    do {                                                                       \
        hipError_t _e = (expr);                                                \
        if (_e != hipSuccess) {                                                \
            char _b[256];                                                      \
            snprintf(_b, sizeof(_b), "%s failed: %s (%s:%d)", #expr,          \
                     hipGetErrorString(_e), __FILE__, __LINE__);               \
            return std::string(_b);                                            \
        }                                                                      \
    } while (0)

namespace {

constexpr unsigned kSplitCoef = 10, kVertCoef = 2;   // kernel.cu:188-191 (K*10, K*2 entries)

struct PointArgs {
    EdgeAoS* edges;
    VertAoS* verts;
    double* cornerA;
    double* cornerB;
    int nv, nf, K, hopCap;
    unsigned ssz, vsz;
    PSplitItem* sp[2];
    PVertItem* vt[2];
    PWinItem* wq[2];
    PPsItem* pq[2];
    Win* stored;
    PointRes* res;
    int f[2];
    double p[2][3];
};

__device__ void make_policy(const PointArgs& a, int run, PlainPolicy& p)
{
    p.E = a.edges; p.Vt = a.verts; p.cA = a.cornerA; p.cB = a.cornerB;
    p.nv_ = a.nv; p.nf_ = a.nf;
    p.ssz = a.ssz; p.vsz = a.vsz;
    p.sp = a.sp[run]; p.vt = a.vt[run]; p.wq = a.wq[run]; p.pq = a.pq[run];
    p.wtail = 1; p.ptail = 1; p.flags = 0;
    p.srcVert = a.nv;            // srcId of a surface point (ICHDevice.cu:596)
}

// Straight-line distances from point q to the vertices of face f, in the
// order of its half-edges (Vector3D::length, Geometry.cuh:95).
__device__ void face_dists(const PointArgs& a, int f, const double* q, int fv[3], double d[3])
{
    for (int j = 0; j < 3; ++j) {
        fv[j] = a.edges[3 * f + j].v0;
        const VertAoS& v = a.verts[fv[j]];
        const double x = q[0] - v.x, y = q[1] - v.y, z = q[2] - v.z;
        d[j] = sqrt(x * x + y * y + z * z);
    }
}

__global__ void __launch_bounds__(128) k_point_ich(PointArgs a)
{
    if ((threadIdx.x & 63) != 0) return;
    const int run = threadIdx.x >> 6;
    PlainPolicy p;
    make_policy(a, run, p);
    p.clear();
    Ich<PlainPolicy> ich(p, -1, a.K);
    ich.srcFace = a.f[run];
    int fv[3];
    double d[3];
    face_dists(a, a.f[run], a.p[run], fv, d);
    ich.initialize_point(a.f[run], d);
    int nStored = 0;
    ich.execute_point(a.stored, run == 0 ? kStoredWin : 0, nStored, run == 0 ? a.f[1] : -1);
    if (run == 1) {
        atomicOr(&a.res->status, p.flags);
        return;
    }
    PointRes r;
    r.srcX = 0.0; r.dstX = 0.0;            // SVGNode X fields start at 0
    face_dists(a, a.f[1], a.p[1], fv, d);
    r.dist = ich.face_path_to(a.f[1], fv, d, a.stored, nStored, r.srcEdge, r.srcX, r.dstEdge, r.dstX, a.hopCap);
    r.lastVert = -1; r.firstVert = -1;
    r.route = 0;
    atomicOr(&a.res->status, p.flags);
    a.res->dist = r.dist; a.res->srcX = r.srcX; a.res->dstX = r.dstX;
    a.res->srcEdge = r.srcEdge; a.res->dstEdge = r.dstEdge;
    a.res->lastVert = r.lastVert; a.res->firstVert = r.firstVert; a.res->route = r.route;
}

constexpr int kJoinThreads = 256;

__global__ void __launch_bounds__(kJoinThreads) k_point_join(PointArgs a, const unsigned long long* gdist,
                                                              const int32_t* gparent)
{
    // min over the p1 table's slots of graph + ICH distance; ties keep the
    // lowest slot (the reference's strict-< scan in slot order)
    __shared__ double sBest[kJoinThreads];
    __shared__ int sSlot[kJoinThreads];
    const PVertItem* vt = a.vt[1];
    double best = DBL_MAX;
    int slot = INT_MAX;
    for (int i = threadIdx.x; i < (int)a.vsz; i += kJoinThreads) {
        const double d1 = vt[i].dist;
        if (d1 == DBL_MAX) continue;
        const double s = __longlong_as_double((long long)gdist[vt[i].index]) + d1;
        if (s < best) { best = s; slot = i; }
    }
    sBest[threadIdx.x] = best;
    sSlot[threadIdx.x] = slot;
    __syncthreads();
    for (int h = kJoinThreads / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            const double ob = sBest[threadIdx.x + h];
            const int os = sSlot[threadIdx.x + h];
            if (ob < sBest[threadIdx.x] || (ob == sBest[threadIdx.x] && os < sSlot[threadIdx.x])) {
                sBest[threadIdx.x] = ob;
                sSlot[threadIdx.x] = os;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    PointRes r = *a.res;             // the direct attempt
    r.route = 1;
    r.dist = sBest[0];
    if (sSlot[0] == INT_MAX || !(sBest[0] < DBL_MAX)) {
        // nothing joins the points (the reference would backtrace vertex
        // 0xFFFFFFFF here)
        r.route = -1; r.dist = DBL_MAX;
        r.srcEdge = r.dstEdge = -1; r.srcX = r.dstX = 0.0;
        r.lastVert = r.firstVert = -1;
        *a.res = r;
        return;
    }
    const int last = vt[sSlot[0]].index;
    r.lastVert = last;
    // p1 side: its source-side key point becomes the destination-side one
    // (SVG.cu:309-318), converted to the twin half-edge
    PlainPolicy p1;
    make_policy(a, 1, p1);
    Ich<PlainPolicy> ich1(p1, -1, a.K);
    ich1.srcFace = a.f[1];
    int nDE, bSE;
    double nDX, bSX;
    nDX = 0.0;
    ich1.path_to(last, nDE, nDX, bSE, bSX, a.hopCap);
    r.srcEdge = bSE; r.srcX = bSX;
    if (nDE == -1) { nDE = r.srcEdge; nDX = r.dstX; }      // reads nextToDstX, as SVG.cu:314 does
    if (nDE != -1 && a.edges[nDE].twin != -1) {
        nDE = a.edges[nDE].twin;
        nDX = a.edges[nDE].len - nDX;
    }
    // first graph vertex: root of the parent chain
    int first = last;
    for (int hops = 0; hops <= a.nv && gparent[first] >= 0 && gparent[first] < a.nv; ++hops)
        first = gparent[first];
    r.firstVert = first;
    // p0 side from the first vertex (SVG.cu:320-351); the p0 tables are the
    // run of k_point_ich (the reference's re-run rebuilds the same state)
    PlainPolicy p0;
    make_policy(a, 0, p0);
    Ich<PlainPolicy> ich0(p0, -1, a.K);
    ich0.srcFace = a.f[0];
    if (p0.vget(first).dist != DBL_MAX) {
        int nSE, aDE;
        double nSX, aDX;
        ich0.path_to(first, nSE, nSX, aDE, aDX, a.hopCap);
        r.dstEdge = aDE; r.dstX = aDX;
        if (nSE == -1) { nSE = r.dstEdge; nSX = r.dstX; }
        r.srcEdge = nSE; r.srcX = nSX;
        r.dstEdge = nDE; r.dstX = nDX;
    }
    *a.res = r;
}

void fill_args(PointWork& w, const DevMesh& m, int f0, const double* p0, int f1, const double* p1, int hopCap,
               PointArgs& a)
{
    a.edges = m.edges; a.verts = m.verts; a.cornerA = m.cornerA; a.cornerB = m.cornerB;
    a.nv = m.nv; a.nf = m.nf; a.K = w.K; a.hopCap = hopCap;
    a.ssz = w.ssz; a.vsz = w.vsz;
    for (int r = 0; r < 2; ++r) {
        a.sp[r] = (PSplitItem*)w.sp[r]; a.vt[r] = (PVertItem*)w.vt[r];
        a.wq[r] = (PWinItem*)w.wq[r]; a.pq[r] = (PPsItem*)w.pq[r];
    }
    a.stored = (Win*)w.stored;
    a.res = w.res;
    a.f[0] = f0; a.f[1] = f1;
    for (int k = 0; k < 3; ++k) {
        a.p[0][k] = p0 ? p0[k] : 0.0;
        a.p[1][k] = p1 ? p1[k] : 0.0;
    }
}

}  // namespace

std::string point_alloc(PointWork& w, int K)
{
    if (w.K == K && w.res) return std::string();
    point_free(w);
    w.K = K;
    w.ssz = (unsigned)K * kSplitCoef;
    w.vsz = (unsigned)K * kVertCoef;
    for (int r = 0; r < 2; ++r) {
        PT_TRY(hipMalloc(&w.sp[r], sizeof(PSplitItem) * (w.ssz + 1)));
        PT_TRY(hipMalloc(&w.vt[r], sizeof(PVertItem) * (w.vsz + 1)));
        PT_TRY(hipMalloc(&w.wq[r], sizeof(PWinItem) * (kHeapCap + 1)));
        PT_TRY(hipMalloc(&w.pq[r], sizeof(PPsItem) * (kHeapCap + 1)));
    }
    PT_TRY(hipMalloc(&w.stored, sizeof(Win) * kStoredWin));
    PT_TRY(hipMalloc(&w.res, sizeof(PointRes)));
    w.hostVt.resize(sizeof(PVertItem) * (w.vsz + 1));
    return std::string();
}

void point_free(PointWork& w)
{
    for (int r = 0; r < 2; ++r) {
        hipFree(w.sp[r]); hipFree(w.vt[r]); hipFree(w.wq[r]); hipFree(w.pq[r]);
        w.sp[r] = w.vt[r] = w.wq[r] = w.pq[r] = nullptr;
    }
    hipFree(w.stored); hipFree(w.res);
    w.stored = nullptr; w.res = nullptr;
    w.K = 0;
}

std::string point_ich(PointWork& w, const DevMesh& m, int f0, const double* p0, int f1, const double* p1,
                      int hopCap, hipStream_t stream, PointRes& out)
{
    if (!m.edges || !m.verts) return "surface-point queries need the reference-layout mesh";
    if (f0 < 0 || f1 < 0 || f0 >= m.nf || f1 >= m.nf) return "face out of range";
    PointArgs a;
    fill_args(w, m, f0, p0, f1, p1, hopCap, a);
    PointRes zero;
    memset(&zero, 0, sizeof(zero));
    PT_TRY(hipMemcpyAsync(w.res, &zero, sizeof(PointRes), hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(k_point_ich, dim3(1), dim3(128), 0, stream, a);
    PT_TRY(hipGetLastError());
    PT_TRY(hipMemcpyAsync(&out, w.res, sizeof(PointRes), hipMemcpyDeviceToHost, stream));
    PT_TRY(hipStreamSynchronize(stream));
    return std::string();
}

std::string point_seeds(PointWork& w, hipStream_t stream, std::vector<int>& src, std::vector<double>& dist)
{
    PT_TRY(hipMemcpyAsync(w.hostVt.data(), w.vt[0], w.hostVt.size(), hipMemcpyDeviceToHost, stream));
    PT_TRY(hipStreamSynchronize(stream));
    const PVertItem* vt = (const PVertItem*)w.hostVt.data();
    src.clear(); dist.clear();
    for (unsigned i = 0; i < w.vsz; ++i) {
        if (vt[i].dist == DBL_MAX) continue;
        src.push_back(vt[i].index);
        dist.push_back(vt[i].dist);
    }
    return std::string();
}

std::string point_join(PointWork& w, const DevMesh& m, const unsigned long long* gdist, const int32_t* gparent,
                       int f0, int f1, int hopCap, hipStream_t stream, PointRes& out)
{
    PointArgs a;
    fill_args(w, m, f0, nullptr, f1, nullptr, hopCap, a);
    hipLaunchKernelGGL(k_point_join, dim3(1), dim3(kJoinThreads), 0, stream, a, gdist, gparent);
    PT_TRY(hipGetLastError());
    PT_TRY(hipMemcpyAsync(&out, w.res, sizeof(PointRes), hipMemcpyDeviceToHost, stream));
    PT_TRY(hipStreamSynchronize(stream));
    return std::string();
}

}  // namespace svg
