// Geodesic distance between two surface points on gfx950 (SVG::SolveSSSD(f0,
// p0, f1, p1), SVG.cu:253-352).
//
// Each query needs two bounded ICH propagations (from p0, keeping the windows
// that cross into f1, and from p1), run in the reference's table layout
// (plain_policy.h); the p0 run then answers directly through f1
// (BuildGeodesicPathTo(face, pos), ICHDevice.cu:143-334).  A single query
// runs each propagation in one block with its heaps in LDS
// (k_point_ich_lds); many queries run one propagation per wave
// (k_point_ich_many).  The reference runs the p0 propagation twice more
// (SVG.cu:294-296, 330-332) only to rebuild state these resident tables
// still hold.  When the direct answer has no source-side key point
// (SVG.cu:284) the graph stage runs on the solver as a join slot (a
// multi-source Dijkstra seeded with the p0 run's vertex distances that stops
// once nothing left can improve the join) and k_point_join takes the minimum
// over the p1 run's vertices of graph distance plus ICH distance (first table
// slot on ties, SVG.cu:299-308) and backtraces both ends (SVG.cu:309-351).
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
// Runs per 64-wide wave in k_point_ich_many: a run is a latency-bound chain,
// so few lanes per wave (little divergence); 4 lanes keep 16K runs resident
// (4 waves per SIMD) where one lane per wave would hold 4K.
constexpr int kRunsPerWave = 4;

// Run r of the batch is query r / 2 from its first point (r even: p0, keeping
// the windows that cross into f1) or its second (r odd: p1).
struct PointArgs {
    EdgeAoS* edges;
    VertAoS* verts;
    double* cornerA;
    double* cornerB;
    int nv, nf, K, hopCap;
    unsigned ssz, vsz;
    PSplitItem* sp;          // per-run tables, run-major
    PVertItem* vt;
    unsigned char* wq;       // per-run window heaps (kWinHeapBytes each)
    PPsItem* pq;
    Win* stored;             // kStoredWin per query
    PointRes* res;           // per query
    const int32_t* f;        // f0, f1 per query
    const double* p;         // p0, p1 (xyz) per query
    int nRuns;
};

// The reference-layout policy with pops that carry only the moving item's
// key through the sift-down: the last item stays in its slot (the sift only
// writes parent slots below it) and is copied once at the end.  Same heap
// contents as PlainPolicy's pops, without an 80-byte item live across the
// loop (which the compiler keeps in scratch memory).
// Window heap of a point run: the reference's 1-based binary heap on the
// same keys with the same tie and capacity rules (PriorityQueue.cuh:51-105),
// holding 16-B (key, record) entries while the windows sit in a pool of
// 56-B records.  A sift moves 16 B per level instead of the 80-B item; the
// pop order is the reference's.
struct PHeapEnt { double key; int idx; int pad; };
struct PRec { double b0, b1, d0, d1, ps; int edge, pseudoSrc, birth, level; };
static_assert(sizeof(PHeapEnt) == 16 && sizeof(PRec) == 56, "record sizes");
constexpr int kPool = kHeapCap + 1;             // records (the heap holds at most kHeapCap)
constexpr size_t kWinHeapBytes = sizeof(PHeapEnt) * (kHeapCap + 1) + sizeof(PRec) * kPool + sizeof(int) * kPool;

struct PointPolicy : PlainPolicy {
    PHeapEnt* wh;            // heap entries 1..kHeapCap
    PRec* wr;                // record pool
    int* wfree;              // free record stack
    int wfreeTop;

    __device__ void bind_window_heap(void* base)
    {
        wh = (PHeapEnt*)base;
        wr = (PRec*)((char*)base + sizeof(PHeapEnt) * (kHeapCap + 1));
        wfree = (int*)((char*)wr + sizeof(PRec) * kPool);
    }
    __device__ void clear()
    {
        PlainPolicy::clear();
        for (int i = 0; i < kPool; ++i) wfree[i] = kPool - 1 - i;
        wfreeTop = kPool;
    }
    __device__ void wtop(Win& w) const
    {
        const PRec& r = wr[wh[1].idx];
        w.edge = r.edge; w.b0 = r.b0; w.b1 = r.b1; w.d0 = r.d0; w.d1 = r.d1; w.ps = r.ps;
        win_finish(w);
        w.twin = E[w.edge].twin;
        w.pseudoSrc = r.pseudoSrc; w.birth = r.birth; w.level = r.level;
    }
    // Both children's entries are read at once, so a level of the sift-down
    // is one round trip; the right child wins ties, as in the reference.
    __device__ void wpop()
    {
        const int freed = wh[1].idx;
        --wtail;
        const PHeapEnt last = wh[wtail];
        int cur = 1;
        for (;;) {
            const int l = 2 * cur, r = l + 1;
            if (l >= wtail) break;
            const PHeapEnt cl = wh[l];
            const PHeapEnt cr = wh[r < wtail ? r : l];
            const bool right = r < wtail && !(cl.key < cr.key);
            const PHeapEnt cm = right ? cr : cl;
            if (!(cm.key < last.key)) break;
            wh[cur] = cm;
            cur = right ? r : l;
        }
        wh[cur] = last;
        wfree[wfreeTop++] = freed;
    }
    __device__ void wpush(const Win& w)
    {
        if (wtail - 1 >= kHeapCap) { flags |= SF_WIN_DROP; return; }
        const int idx = wfree[--wfreeTop];
        PRec& rec = wr[idx];
        rec.b0 = w.b0; rec.b1 = w.b1; rec.d0 = w.d0; rec.d1 = w.d1; rec.ps = w.ps;
        rec.edge = w.edge; rec.pseudoSrc = w.pseudoSrc; rec.birth = w.birth; rec.level = w.level;
        ++wtail;
        int cur = wtail - 1, pp = cur / 2;
        const double key = w.minDist;
        while (pp > 0 && wh[pp].key > key) { wh[cur] = wh[pp]; cur = pp; pp = cur / 2; }
        PHeapEnt e; e.key = key; e.idx = idx; e.pad = 0;
        wh[cur] = e;
    }
    __device__ void ppop()
    {
        --ptail;
        const double key = pq[ptail].key;
        int cur = 1, l = 2, r = 3, mc = -1;
        if (r < ptail) mc = pq[l].key < pq[r].key ? l : r;
        else if (l < ptail) mc = l;
        while (mc != -1 && pq[mc].key < key) {
            pq[cur] = pq[mc];
            cur = mc; l = cur * 2; r = cur * 2 + 1;
            if (r < ptail) mc = pq[l].key < pq[r].key ? l : r;
            else if (l < ptail) mc = l;
            else mc = -1;
        }
        if (cur != ptail) pq[cur] = pq[ptail];
    }
};

__device__ void make_policy(const PointArgs& a, int run, PlainPolicy& p)
{
    p.E = a.edges; p.Vt = a.verts; p.cA = a.cornerA; p.cB = a.cornerB;
    p.nv_ = a.nv; p.nf_ = a.nf;
    p.ssz = a.ssz; p.vsz = a.vsz;
    p.sp = a.sp + (size_t)run * (a.ssz + 1);
    p.vt = a.vt + (size_t)run * (a.vsz + 1);
    p.wq = nullptr;              // the window heap is PointPolicy's (bind_window_heap)
    p.pq = a.pq + (size_t)run * (kHeapCap + 1);
    p.wtail = 1; p.ptail = 1; p.flags = 0;
    p.srcVert = a.nv;            // srcId of a surface point (ICHDevice.cu:596)
}

__device__ void make_policy(const PointArgs& a, int run, PointPolicy& p)
{
    make_policy(a, run, static_cast<PlainPolicy&>(p));
    p.bind_window_heap((char*)a.wq + (size_t)run * kWinHeapBytes);
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

// One ICH run from its point; the p0 run then answers directly through f1.
__device__ void point_run(const PointArgs& a, int run, PointPolicy& p)
{
    const int q = run >> 1, side = run & 1;
    const int f = a.f[2 * q + side];
    p.clear();
    Ich<PointPolicy> ich(p, -1, a.K);
    ich.srcFace = f;
    int fv[3];
    double d[3];
    face_dists(a, f, a.p + 6 * q + 3 * side, fv, d);
    ich.initialize_point(f, d);
    int nStored = 0;
    Win* stored = a.stored + (size_t)q * kStoredWin;
    ich.execute_point(stored, side == 0 ? kStoredWin : 0, nStored, side == 0 ? a.f[2 * q + 1] : -1);
    PointRes* res = a.res + q;
    if (side == 0) {
        PointRes r;
        r.srcX = 0.0; r.dstX = 0.0;            // SVGNode X fields start at 0
        face_dists(a, a.f[2 * q + 1], a.p + 6 * q + 3, fv, d);
        r.dist = ich.face_path_to(a.f[2 * q + 1], fv, d, stored, nStored, r.srcEdge, r.srcX, r.dstEdge, r.dstX,
                                  a.hopCap);
        res->dist = r.dist; res->srcX = r.srcX; res->dstX = r.dstX;
        res->srcEdge = r.srcEdge; res->dstEdge = r.dstEdge;
        res->lastVert = -1; res->firstVert = -1; res->route = 0;
    }
    atomicOr(&res->status, p.flags | (ich.stalled ? (side == 0 ? SF_STALL_P0 : SF_STALL_P1) : 0u));
}

// LDS bytes of a run's window heap, pseudo-source heap and vertex table.
inline size_t lds_bytes(unsigned vsz)
{
    return kWinHeapBytes + sizeof(PPsItem) * (kHeapCap + 1) + sizeof(PVertItem) * (vsz + 1);
}

// One block per run, for a single query.  A run is one sequential
// propagation, a chain of dependent heap and table accesses: its two heaps
// and its vertex table live in the block's LDS (160 KB on gfx950; the split
// table stays in global memory) and the vertex table is written back for the
// join.
__global__ void __launch_bounds__(64) k_point_ich_lds(PointArgs a)
{
    extern __shared__ double ldsBuf[];
    const int run = blockIdx.x;
    PointPolicy p;
    make_policy(a, run, p);
    PVertItem* dst = p.vt;
    unsigned char* b = (unsigned char*)ldsBuf;
    p.bind_window_heap(b);
    p.pq = (PPsItem*)(b + kWinHeapBytes);
    p.vt = (PVertItem*)(b + kWinHeapBytes + sizeof(PPsItem) * (kHeapCap + 1));
    if (threadIdx.x == 0) point_run(a, run, p);
    __syncthreads();
    for (unsigned i = threadIdx.x; i < a.vsz + 1; i += blockDim.x) dst[i] = p.vt[i];
}

// Many queries: one run per lane of the first kRunsPerWave lanes of each
// wave, tables in global memory (the plain build kernel's layout), so the
// device holds thousands of runs at once.
// Three waves per SIMD (168 registers, no VGPR spills) beat two (the
// unconstrained 172 registers): 1,360 against 1,411 ms for 8,192 queries.
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3, 3))) k_point_ich_many(PointArgs a)
{
    if ((int)threadIdx.x >= kRunsPerWave) return;
    const int total = gridDim.x * kRunsPerWave;
    for (int run = blockIdx.x * kRunsPerWave + threadIdx.x; run < a.nRuns; run += total) {
        PointPolicy p;
        make_policy(a, run, p);
        point_run(a, run, p);
    }
}

constexpr int kJoinThreads = 256;

// One block per finished graph stage: jobs[2b] = query, jobs[2b+1] = its
// solver slot (graph arrays at slot * V).
__global__ void __launch_bounds__(kJoinThreads) k_point_join(PointArgs a, const int32_t* jobs,
                                                              const unsigned long long* gdistAll,
                                                              const int32_t* gparentAll)
{
    const int q = jobs[2 * blockIdx.x];
    const size_t sb = (size_t)jobs[2 * blockIdx.x + 1] * (size_t)a.nv;
    const unsigned long long* gdist = gdistAll + sb;
    const int32_t* gparent = gparentAll + sb;
    // min over the p1 table's slots of graph + ICH distance; ties keep the
    // lowest slot (the reference's strict-< scan in slot order)
    __shared__ double sBest[kJoinThreads];
    __shared__ int sSlot[kJoinThreads];
    const PVertItem* vt = a.vt + (size_t)(2 * q + 1) * (a.vsz + 1);
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
    PointRes r = a.res[q];           // the direct attempt
    r.route = 1;
    r.dist = sBest[0];
    if (sSlot[0] == INT_MAX || !(sBest[0] < DBL_MAX)) {
        // nothing joins the points (the reference would backtrace vertex
        // 0xFFFFFFFF here)
        r.route = -1; r.dist = DBL_MAX;
        r.srcEdge = r.dstEdge = -1; r.srcX = r.dstX = 0.0;
        r.lastVert = r.firstVert = -1;
        a.res[q] = r;
        return;
    }
    const int last = vt[sSlot[0]].index;
    r.lastVert = last;
    // p1 side: its source-side key point becomes the destination-side one
    // (SVG.cu:309-318), converted to the twin half-edge
    PlainPolicy p1;
    make_policy(a, 2 * q + 1, p1);
    Ich<PlainPolicy> ich1(p1, -1, a.K);
    ich1.srcFace = a.f[2 * q + 1];
    int nDE, bSE;
    double nDX, bSX;
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
    // run above (the reference's re-run rebuilds the same state)
    PlainPolicy p0;
    make_policy(a, 2 * q, p0);
    Ich<PlainPolicy> ich0(p0, -1, a.K);
    ich0.srcFace = a.f[2 * q];
    if (p0.vget(first).dist != DBL_MAX) {
        int nSE, aDE;
        double nSX, aDX;
        ich0.path_to(first, nSE, nSX, aDE, aDX, a.hopCap);
        r.dstEdge = aDE; r.dstX = aDX;
        if (nSE == -1) { nSE = r.dstEdge; nSX = r.dstX; }
        r.srcEdge = nSE; r.srcX = nSX;
        r.dstEdge = nDE; r.dstX = nDX;
    }
    a.res[q] = r;
}

PointArgs make_args(PointWork& w, const DevMesh& m, int hopCap, int nRuns)
{
    PointArgs a;
    a.edges = m.edges; a.verts = m.verts; a.cornerA = m.cornerA; a.cornerB = m.cornerB;
    a.nv = m.nv; a.nf = m.nf; a.K = w.K; a.hopCap = hopCap;
    a.ssz = w.ssz; a.vsz = w.vsz;
    a.sp = (PSplitItem*)w.sp; a.vt = (PVertItem*)w.vt; a.wq = (unsigned char*)w.wq; a.pq = (PPsItem*)w.pq;
    a.stored = (Win*)w.stored;
    a.res = w.res;
    a.f = w.fDev; a.p = w.pDev;
    a.nRuns = nRuns;
    return a;
}

}  // namespace

size_t point_bytes_per_query(int K)
{
    const size_t ssz = (size_t)K * kSplitCoef, vsz = (size_t)K * kVertCoef;
    return 2 * (sizeof(PSplitItem) * (ssz + 1) + sizeof(PVertItem) * (vsz + 1) + kWinHeapBytes +
                sizeof(PPsItem) * (kHeapCap + 1)) +
           sizeof(Win) * kStoredWin + sizeof(PointRes) + 2 * sizeof(int32_t) + 6 * sizeof(double);
}

std::string point_alloc(PointWork& w, int K, int cap)
{
    if (w.K == K && w.cap >= cap && w.res) return std::string();
    point_free(w);
    w.K = K;
    w.cap = cap;
    w.ssz = (unsigned)K * kSplitCoef;
    w.vsz = (unsigned)K * kVertCoef;
    const size_t runs = 2 * (size_t)cap;
    PT_TRY(hipMalloc(&w.sp, sizeof(PSplitItem) * (w.ssz + 1) * runs));
    PT_TRY(hipMalloc(&w.vt, sizeof(PVertItem) * (w.vsz + 1) * runs));
    PT_TRY(hipMalloc(&w.wq, kWinHeapBytes * runs));
    PT_TRY(hipMalloc(&w.pq, sizeof(PPsItem) * (kHeapCap + 1) * runs));
    PT_TRY(hipMalloc(&w.stored, sizeof(Win) * kStoredWin * (size_t)cap));
    PT_TRY(hipMalloc(&w.res, sizeof(PointRes) * (size_t)cap));
    PT_TRY(hipMalloc(&w.fDev, sizeof(int32_t) * 2 * (size_t)cap));
    PT_TRY(hipMalloc(&w.pDev, sizeof(double) * 6 * (size_t)cap));
    PT_TRY(hipMalloc(&w.jobsDev, sizeof(int32_t) * 2 * (size_t)cap));
    PT_TRY(hipHostMalloc(&w.jobsHost, sizeof(int32_t) * 2 * (size_t)cap, hipHostMallocDefault));
    w.hostVt.resize(sizeof(PVertItem) * (w.vsz + 1));
    return std::string();
}

void point_free(PointWork& w)
{
    hipFree(w.sp); hipFree(w.vt); hipFree(w.wq); hipFree(w.pq);
    hipFree(w.stored); hipFree(w.res); hipFree(w.fDev); hipFree(w.pDev); hipFree(w.jobsDev);
    if (w.jobsHost) hipHostFree(w.jobsHost);
    w.sp = w.vt = w.wq = w.pq = w.stored = nullptr;
    w.res = nullptr; w.fDev = nullptr; w.pDev = nullptr; w.jobsDev = nullptr; w.jobsHost = nullptr;
    w.K = 0; w.cap = 0;
}

std::string point_ich(PointWork& w, const DevMesh& m, int n, const int32_t* faces, const double* pts, int hopCap,
                      hipStream_t stream, PointRes* out)
{
    if (!m.edges || !m.verts) return "surface-point queries need the reference-layout mesh";
    if (n < 1 || n > w.cap) return "query count out of range";
    for (int i = 0; i < 2 * n; ++i)
        if (faces[i] < 0 || faces[i] >= m.nf) return "face out of range";
    PT_TRY(hipMemcpyAsync(w.fDev, faces, sizeof(int32_t) * 2 * (size_t)n, hipMemcpyHostToDevice, stream));
    PT_TRY(hipMemcpyAsync(w.pDev, pts, sizeof(double) * 6 * (size_t)n, hipMemcpyHostToDevice, stream));
    PT_TRY(hipMemsetAsync(w.res, 0, sizeof(PointRes) * (size_t)n, stream));
    PointArgs a = make_args(w, m, hopCap, 2 * n);
    if (n == 1 && w.ldsOk < 0) {
        // LDS-resident runs when the heaps and the vertex table fit a block
        int dev = 0, maxLds = 0;
        PT_TRY(hipGetDevice(&dev));
        PT_TRY(hipDeviceGetAttribute(&maxLds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        w.ldsOk = lds_bytes(w.vsz) <= (size_t)maxLds ? 1 : 0;
        if (w.ldsOk)
            PT_TRY(hipFuncSetAttribute((const void*)k_point_ich_lds, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds_bytes(w.vsz)));
    }
    if (n == 1 && w.ldsOk == 1) {
        hipLaunchKernelGGL(k_point_ich_lds, dim3(2), dim3(64), lds_bytes(w.vsz), stream, a);
    } else {
        const int blocks = (2 * n + kRunsPerWave - 1) / kRunsPerWave;
        hipLaunchKernelGGL(k_point_ich_many, dim3(blocks), dim3(64), 0, stream, a);
    }
    PT_TRY(hipGetLastError());
    PT_TRY(hipMemcpyAsync(out, w.res, sizeof(PointRes) * (size_t)n, hipMemcpyDeviceToHost, stream));
    PT_TRY(hipStreamSynchronize(stream));
    return std::string();
}

std::string point_seeds(PointWork& w, int q, int side, hipStream_t stream, std::vector<int>& src,
                        std::vector<double>& dist)
{
    const PVertItem* dvt = (const PVertItem*)w.vt + (size_t)(2 * q + side) * (w.vsz + 1);
    PT_TRY(hipMemcpyAsync(w.hostVt.data(), dvt, w.hostVt.size(), hipMemcpyDeviceToHost, stream));
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

std::string point_join(PointWork& w, const DevMesh& m, const std::vector<int32_t>& jobs,
                       const unsigned long long* gdist, const int32_t* gparent, int hopCap, hipStream_t stream)
{
    const int n = (int)(jobs.size() / 2);
    if (n == 0) return std::string();
    if (n > w.cap) return "too many joins";
    // the job list travels through pinned memory; the previous batch's
    // copy has completed once its join kernel has (same stream)
    PT_TRY(hipStreamSynchronize(stream));
    memcpy(w.jobsHost, jobs.data(), sizeof(int32_t) * jobs.size());
    PT_TRY(hipMemcpyAsync(w.jobsDev, w.jobsHost, sizeof(int32_t) * jobs.size(), hipMemcpyHostToDevice, stream));
    PointArgs a = make_args(w, m, hopCap, 0);
    hipLaunchKernelGGL(k_point_join, dim3(n), dim3(kJoinThreads), 0, stream, a, w.jobsDev, gdist, gparent);
    PT_TRY(hipGetLastError());
    return std::string();
}

std::string point_results(PointWork& w, int n, hipStream_t stream, PointRes* out)
{
    PT_TRY(hipMemcpyAsync(out, w.res, sizeof(PointRes) * (size_t)n, hipMemcpyDeviceToHost, stream));
    PT_TRY(hipStreamSynchronize(stream));
    return std::string();
}

}  // namespace svg
