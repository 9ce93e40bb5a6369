// SVG construction on gfx950: the plain one-source-per-thread form and the
// batched persistent form.  Both run Ich<> (ich_engine.h); they differ only in
// how the per-source state is stored and scheduled.
//
// plain   (SVG.cu:9-96 structure): reference records (80-B heap items, the
//         MyHashTable with its hash and overflow slot), per-thread scratch,
//         Clear() of every table slot per source, grid-stride over sources,
//         O(V) neighbour scan.  Exact even where the reference's tables
//         overflow; used to recompute sources the batched form flags.
// batched persistent lanes take sources from an atomic counter; a source's
//         vertex / split tables are compact power-of-two open-addressing
//         tables that stay resident until extraction; the window heap keeps
//         16-B (key, record) entries with 64-B records in a slot pool and the
//         root in registers.  Extraction runs one wave per source: the
//         reached list is sorted in LDS, 64 backtraces run at once, and a
//         ballot + popcount keeps the first K qualifying vertices in id order
//         (SVG.cu:44-76 semantics without the O(V) scan).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <system_error>
#include <thread>
#include <vector>

#include "svg_acos.h"
#include "ich_engine.h"
#include "svg_build.h"
#include "plain_policy.h"
#include "source_order.h"
#include "host_parallel.h"

namespace svg {

#define HIP_TRY(expr)                                                          \
    do {                                                                       \
        hipError_t _e = (expr);                                                \
        if (_e != hipSuccess) {                                                \
            char _b[256];                                                      \
            snprintf(_b, sizeof(_b), "%s failed: %s (%s:%d)", #expr,          \
                     hipGetErrorString(_e), __FILE__, __LINE__);               \
            return std::string(_b);                                            \
        }                                                                      \
    } while (0)

// ---------------------------------------------------------------------------
// mesh upload
// ---------------------------------------------------------------------------
std::string upload_mesh(const HostMesh& h, DevMesh& d, bool aos)
{
    d.nv = h.nv; d.nf = h.nf; d.ne = h.ne;
    if (!d.face) {
        std::vector<FaceRec> fr((size_t)h.nf);
        parallel_for(h.nf, [&](int f0, int f1) {
            for (int f = f0; f < f1; ++f) {
                FaceRec& r = fr[f];
                for (int j = 0; j < 3; ++j) {
                    r.l[j] = h.len[3 * f + j];
                    r.v[j] = h.tri[3 * f + j];
                    r.tw[j] = h.twin[3 * f + j];
                }
                r.pad[0] = r.pad[1] = r.pad[2] = r.pad[3] = 0;
            }
        });
        std::vector<double> pos(4 * (size_t)h.nv);
        parallel_for(h.nv, [&](int v0, int v1) {
            for (int v = v0; v < v1; ++v) {
                pos[4 * v] = h.xyz[3 * v]; pos[4 * v + 1] = h.xyz[3 * v + 1];
                pos[4 * v + 2] = h.xyz[3 * v + 2]; pos[4 * v + 3] = 0.0;
            }
        });
        HIP_TRY(hipMalloc(&d.face, sizeof(FaceRec) * (size_t)h.nf));
        HIP_TRY(hipMalloc(&d.cornerA, sizeof(double) * (size_t)h.ne));
        HIP_TRY(hipMalloc(&d.cornerB, sizeof(double) * (size_t)h.ne));
        HIP_TRY(hipMalloc(&d.firstEdge, sizeof(int32_t) * (size_t)h.nv));
        HIP_TRY(hipMalloc(&d.flags, (size_t)h.nv));
        HIP_TRY(hipMalloc(&d.pos, sizeof(double) * 4 * (size_t)h.nv));
        HIP_TRY(hipMemcpy(d.face, fr.data(), sizeof(FaceRec) * (size_t)h.nf, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d.cornerA, h.cornerA.data(), sizeof(double) * (size_t)h.ne, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d.cornerB, h.cornerB.data(), sizeof(double) * (size_t)h.ne, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d.firstEdge, h.firstEdge.data(), sizeof(int32_t) * (size_t)h.nv, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d.flags, h.flags.data(), (size_t)h.nv, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d.pos, pos.data(), sizeof(double) * 4 * (size_t)h.nv, hipMemcpyHostToDevice));
        // fans (only when every half-edge lies on the fan its origin's
        // firstEdge starts: a manifold mesh).  A vertex's fan then holds all
        // of its outgoing half-edges, so the fans start at the prefix sums of
        // the out-degrees and are filled vertex by vertex in parallel; a walk
        // that leaves some outgoing half-edge out marks the mesh non-manifold.
        std::vector<FanRec> fan((size_t)h.ne);
        std::vector<int32_t> fidx((size_t)h.ne, -1), fbase((size_t)h.nv + 1, 0);
        for (int e = 0; e < h.ne; ++e) fbase[h.tri[e] + 1]++;
        for (int v = 0; v < h.nv; ++v) fbase[v + 1] += fbase[v];
        std::atomic<bool> manifold{true};
        auto nxt = [](int e) { return (e % 3 == 2) ? e - 2 : e + 1; };
        auto prv = [](int e) { return (e % 3 == 0) ? e + 2 : e - 1; };
        parallel_for(h.nv, [&](int va, int vb) {
            for (int v = va; v < vb && manifold.load(std::memory_order_relaxed); ++v) {
                int fp = fbase[v];
                const int end = fbase[v + 1];
                const int start = h.firstEdge[v];
                int e = start;
                while (e != -1) {
                    // origin first: fidx of another vertex's half-edge belongs to another thread
                    if (h.tri[e] != v || fp >= end || fidx[e] != -1) { manifold = false; break; }
                    fidx[e] = fp;
                    FanRec& r = fan[fp++];
                    const int tw = h.twin[e];
                    r.cA = h.cornerA[e];
                    r.cB = tw != -1 ? h.cornerB[tw] : 0.0;
                    r.lo = h.len[e]; r.lnx = h.len[nxt(e)]; r.lpv = h.len[prv(e)];
                    r.o = e; r.twn = h.twin[nxt(e)];
                    r.nb = h.tri[nxt(e)];
                    r.tw = tw; r.nbFlag = h.flags[r.nb]; r.pad = 0;
                    e = tw != -1 ? nxt(tw) : -1;
                    if (e == start) break;
                }
                if (fp != end) manifold = false;
            }
        });
        if (manifold) {
            HIP_TRY(hipMalloc(&d.fan, sizeof(FanRec) * (size_t)h.ne));
            HIP_TRY(hipMalloc(&d.fanIdx, sizeof(int32_t) * (size_t)h.ne));
            HIP_TRY(hipMalloc(&d.fanBase, sizeof(int32_t) * ((size_t)h.nv + 1)));
            HIP_TRY(hipMemcpy(d.fan, fan.data(), sizeof(FanRec) * (size_t)h.ne, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(d.fanIdx, fidx.data(), sizeof(int32_t) * (size_t)h.ne, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(d.fanBase, fbase.data(), sizeof(int32_t) * ((size_t)h.nv + 1), hipMemcpyHostToDevice));
        }
    }
    if (aos && !d.edges) {
        std::vector<EdgeAoS> ea((size_t)h.ne);
        for (int e = 0; e < h.ne; ++e) {
            EdgeAoS& x = ea[e];
            int f = e / 3, j = e % 3;
            x.v0 = h.tri[e];
            x.v1 = h.tri[3 * f + (j + 1) % 3];
            x.twin = h.twin[e];
            x.next = 3 * f + (j + 1) % 3;
            x.prev = 3 * f + (j + 2) % 3;
            x.face = f;
            x.len = h.len[e];
        }
        std::vector<VertAoS> va((size_t)h.nv);
        for (int v = 0; v < h.nv; ++v) {
            va[v].x = h.xyz[3 * v]; va[v].y = h.xyz[3 * v + 1]; va[v].z = h.xyz[3 * v + 2];
            va[v].angle = h.angle[v];
            va[v].firstEdge = h.firstEdge[v];
            va[v].flags = h.flags[v];
        }
        HIP_TRY(hipMalloc(&d.edges, sizeof(EdgeAoS) * (size_t)h.ne));
        HIP_TRY(hipMalloc(&d.verts, sizeof(VertAoS) * (size_t)h.nv));
        HIP_TRY(hipMemcpy(d.edges, ea.data(), sizeof(EdgeAoS) * (size_t)h.ne, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d.verts, va.data(), sizeof(VertAoS) * (size_t)h.nv, hipMemcpyHostToDevice));
    }
    return std::string();
}

void free_dev_mesh(DevMesh& d)
{
    hipFree(d.face); hipFree(d.cornerA); hipFree(d.cornerB); hipFree(d.firstEdge);
    hipFree(d.flags); hipFree(d.pos); hipFree(d.edges); hipFree(d.verts);
    hipFree(d.fan); hipFree(d.fanIdx); hipFree(d.fanBase);
    d = DevMesh();
}

std::string alloc_graph(int V, int K, DevGraph& g)
{
    g.V = V; g.K = K;
    const size_t n = (size_t)V * (size_t)K;
    HIP_TRY(hipMalloc(&g.tail, sizeof(int32_t) * (size_t)V));
    HIP_TRY(hipMalloc(&g.adj, sizeof(int32_t) * n));
    HIP_TRY(hipMalloc(&g.w, sizeof(double) * n));
    HIP_TRY(hipMalloc(&g.srcEdge, sizeof(int32_t) * n));
    HIP_TRY(hipMalloc(&g.dstEdge, sizeof(int32_t) * n));
    HIP_TRY(hipMalloc(&g.srcX, sizeof(double) * n));
    HIP_TRY(hipMalloc(&g.dstX, sizeof(double) * n));
    HIP_TRY(hipMemset(g.tail, 0, sizeof(int32_t) * (size_t)V));
    return std::string();
}

void free_graph(DevGraph& g)
{
    hipFree(g.tail); hipFree(g.adj); hipFree(g.w); hipFree(g.srcEdge); hipFree(g.dstEdge);
    hipFree(g.srcX); hipFree(g.dstX);
    g = DevGraph();
}

// ---------------------------------------------------------------------------
// plain form (records and storage policy in plain_policy.h)
// ---------------------------------------------------------------------------
struct PlainArgs {
    DevMesh m;
    DevGraph g;
    int K;
    unsigned ssz, vsz;
    PWinItem* wq;
    PPsItem* pq;
    PSplitItem* sp;
    PVertItem* vt;
    const int32_t* srcList;     // null: all vertices
    int nsrc;
    uint32_t* status;           // per vertex
    int hopCap;
};

__global__ void __launch_bounds__(256) plain_kernel(PlainArgs a)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int total = gridDim.x * blockDim.x;
    PlainPolicy p;
    p.E = a.m.edges; p.Vt = a.m.verts; p.cA = a.m.cornerA; p.cB = a.m.cornerB;
    p.nv_ = a.m.nv; p.nf_ = a.m.nf;
    p.ssz = a.ssz; p.vsz = a.vsz;
    p.sp = a.sp + (size_t)idx * (a.ssz + 1);
    p.vt = a.vt + (size_t)idx * (a.vsz + 1);
    p.wq = a.wq + (size_t)idx * (kHeapCap + 1);
    p.pq = a.pq + (size_t)idx * (kHeapCap + 1);
    const int K = a.K;
    for (int s = idx; s < a.nsrc; s += total) {
        const int i = a.srcList ? a.srcList[s] : s;
        p.srcVert = i;
        p.clear();
        Ich<PlainPolicy> ich(p, i, K);
        ich.execute();
        int tail = 0;
        const size_t row = (size_t)i * (size_t)K;
        for (int j = 0; j < a.m.nv; ++j) {
            const VInfo vj = p.vget(j);
            if (vj.dist == DBL_MAX) continue;
            if (j == i) continue;
            int se, de;
            double sx, dx;
            const bool pass = ich.path_to(j, se, sx, de, dx, a.hopCap);
            if (pass) continue;
            if (se == -1) { se = de; sx = dx; }
            a.g.adj[row + tail] = j;
            a.g.w[row + tail] = vj.dist;
            a.g.srcEdge[row + tail] = se; a.g.srcX[row + tail] = sx;
            a.g.dstEdge[row + tail] = de; a.g.dstX[row + tail] = dx;
            ++tail;
            if (tail == K) break;
        }
        a.g.tail[i] = tail;
        if (a.status) a.status[i] = p.flags;
    }
}

static int hop_cap(const HostMesh& h) { return 4 * (h.nf + h.nv) + 16; }

static double elapsed_ms(hipEvent_t a, hipEvent_t b)
{
    float ms = 0.f;
    hipEventElapsedTime(&ms, a, b);
    return (double)ms;
}

std::string build_plain(const HostMesh& h, const DevMesh& d, const BuildConfig& c, DevGraph& g,
                        BuildStats& st, const int32_t* srcList, int nList)
{
    if (!d.edges) return "plain build needs the reference-layout mesh (upload with aos)";
    const int K = c.K;
    const int blocks = c.plainBlocks > 0 ? c.plainBlocks : 64;
    const int threads = c.plainThreads > 0 ? c.plainThreads : 128;
    const long long nthr = (long long)blocks * threads;
    const unsigned ssz = (unsigned)K * c.splitCoef, vsz = (unsigned)K * c.vertCoef;
    auto t0 = std::chrono::steady_clock::now();
    PlainArgs a;
    a.m = d; a.g = g; a.K = K; a.ssz = ssz; a.vsz = vsz;
    a.wq = nullptr; a.pq = nullptr; a.sp = nullptr; a.vt = nullptr;
    a.srcList = nullptr; a.nsrc = nList > 0 ? nList : h.nv; a.status = nullptr;
    a.hopCap = hop_cap(h);
    int32_t* dList = nullptr;
    std::string err;
    do {
        if (hipMalloc(&a.wq, sizeof(PWinItem) * (size_t)(kHeapCap + 1) * nthr) != hipSuccess ||
            hipMalloc(&a.pq, sizeof(PPsItem) * (size_t)(kHeapCap + 1) * nthr) != hipSuccess ||
            hipMalloc(&a.sp, sizeof(PSplitItem) * (size_t)(ssz + 1) * nthr) != hipSuccess ||
            hipMalloc(&a.vt, sizeof(PVertItem) * (size_t)(vsz + 1) * nthr) != hipSuccess ||
            hipMalloc(&a.status, sizeof(uint32_t) * (size_t)h.nv) != hipSuccess) {
            err = "plain build: out of device memory for per-thread scratch";
            break;
        }
        if (hipMemset(a.status, 0, sizeof(uint32_t) * (size_t)h.nv) != hipSuccess) {
            err = "plain build: clearing the status array failed";
            break;
        }
        if (nList > 0) {
            if (hipMalloc(&dList, sizeof(int32_t) * (size_t)nList) != hipSuccess) {
                err = "plain build: out of device memory for the source list";
                break;
            }
            if (hipMemcpy(dList, srcList, sizeof(int32_t) * (size_t)nList, hipMemcpyHostToDevice) != hipSuccess) {
                err = "plain build: source list upload failed";
                break;
            }
            a.srcList = dList;
        }
        st.allocSeconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        hipEvent_t e0, e1;
        hipEventCreate(&e0); hipEventCreate(&e1);
        hipEventRecord(e0, 0);
        hipLaunchKernelGGL(plain_kernel, dim3(blocks), dim3(threads), 0, 0, a);
        hipEventRecord(e1, 0);
        hipError_t ke = hipEventSynchronize(e1);
        if (ke == hipSuccess) ke = hipGetLastError();
        if (ke != hipSuccess) { err = std::string("plain kernel: ") + hipGetErrorString(ke); }
        const double ms = elapsed_ms(e0, e1);
        if (nList > 0) st.redoSeconds += ms * 1e-3; else st.propSeconds += ms * 1e-3;
        st.seconds += ms * 1e-3;
        hipEventDestroy(e0); hipEventDestroy(e1);
        if (err.empty()) {
            std::vector<uint32_t> stv((size_t)h.nv);
            const hipError_t ce = hipMemcpy(stv.data(), a.status, sizeof(uint32_t) * (size_t)h.nv, hipMemcpyDeviceToHost);
            if (ce != hipSuccess) { err = std::string("plain build: status to host: ") + hipGetErrorString(ce); break; }
            if (st.status.size() != (size_t)h.nv) st.status.assign((size_t)h.nv, 0);
            for (int k = 0; k < a.nsrc; ++k) {
                const int v = nList > 0 ? srcList[k] : k;
                st.status[v] = stv[v] | (nList > 0 ? SF_REDONE : 0);
            }
        }
    } while (0);
    hipFree(a.wq); hipFree(a.pq); hipFree(a.sp); hipFree(a.vt); hipFree(a.status); hipFree(dList);
    st.lanes = (int)nthr;
    return err;
}

// ---------------------------------------------------------------------------
// batched form
// ---------------------------------------------------------------------------
struct alignas(16) BSlotV { int32_t key; uint32_t eb; double dist; };     // 16 B, one load
struct alignas(16) BSlotS { int32_t key; int32_t ps; double dist; double x; double pad; };  // 32 B, two loads
struct BWinRec { double b0, b1, ps, sx, sy; int32_t edge, pseudoSrc, lvlb, twin; double pad; };  // 64 B
static_assert(sizeof(BWinRec) == 64, "window records fill one 64-B half-line");
struct BPsRec { int32_t vert, pseudoSrc, lvlb, pad; };        // 16 B pseudo-window record
struct BHeapEnt { double key; int32_t idx; int32_t pad; };     // 16 B

constexpr uint32_t kEnterNone = 0xFFFFFFu;

// Heap placement: logical binary-heap index i (1-based) keeps the
// reference's semantics; only where entries live changes.  Levels 0-4
// (i < 32) sit in LDS, lane-major so a wave's lanes hit distinct banks (8-B
// keys and 16-bit record indices in separate arrays); levels 5-7 sit in HBM
// as 32 blocks of 128 B, each holding a level-5 node with its children and
// grandchildren (slots 1..7 of eight 16-B entries), and levels 8-9 as 256
// blocks of 64 B, each a level-8 node with its two children (slots 0..2).
// A sift-down below LDS reads two adjacent blocks per step: levels 5-7,
// then levels 8-9.
constexpr int kHeapLds = 32;                          // logical 1..31 in LDS
constexpr int kMidBlocks = 32;                        // levels 5-7, 8 slots each
constexpr int kDeepBase = kMidBlocks * 8;             // levels 8-9: 4 slots per block after them
constexpr int kHeapSlotsMid = kDeepBase;              // storage below 256 items (max index 255)
constexpr int kHeapSlotsFull = kDeepBase + 256 * 4;   // storage up to the reference's 1023 items

__device__ __forceinline__ int bslot(int i)           // i >= 32
{
    const int d = 31 - __clz(i);                      // 5..9
    if (d <= 7) {
        const int r = d - 5;
        return ((i >> r) - 32) * 8 + ((1 << r) | (i & ((1 << r) - 1)));
    }
    const int r = d - 8;                              // 0: the level-8 node, 1: one of its children
    return kDeepBase + ((i >> r) - 256) * 4 + (r ? 1 + (i & 1) : 0);
}

// Multiplicative hashing (Knuth): the slot comes from bits 11 and up of
// k * 2654435761.  (The low bits of the product are a permutation of the
// low bits of k, so runs of ids, as a mesh region's vertices and edges
// are, would land on runs of slots and merge into long probe clusters.)
__device__ __forceinline__ unsigned hmix(unsigned k) { return (k * 2654435761u) >> 11; }
// Vertex-table slot: each aligned group of eight consecutive ids keeps one
// 128-B line of slots (vertices of one face and of neighbouring faces share
// lines), and the groups are scattered by the hash.
__device__ __forceinline__ unsigned vslot(unsigned v) { return (hmix(v >> 3) << 3) | (v & 7u); }

struct BatchedPolicy {
    static constexpr bool kFanRecords = true;
    const FaceRec* __restrict__ F;
    const FanRec* __restrict__ fanR;
    const int32_t* __restrict__ fanIdx;
    const int32_t* __restrict__ fanBase;
    const double* __restrict__ cA;
    const double* __restrict__ cB;
    const int32_t* __restrict__ firstE;
    const uint8_t* __restrict__ vflags;
    int nv_, nf_;
    int src;
    // tables of the current source
    BSlotV* vt; int32_t* rl; unsigned vmask; int vcount; int vrefcap; int vlimit;
    BSlotS* st; unsigned smask; int scount; int srefcap; int slimit;
    // window heap
    BHeapEnt* wh; BWinRec* wr; uint16_t* wfree;
    int wcap, pcap;                  // queue storage per lane (<= kHeapCap); wmax, pmax: peak lengths
    int wmax, pmax;
    double* lk; uint16_t* li; int lstride;   // LDS part of the heap (logical 1..31), lane-major
    int wtail, wftop, wfresh, wcache, wcache2;
    double rootKey; int rootIdx;
    BWinRec nrec; int nrecIdx;       // record of the root, loaded right after a pop
    // The last entry of each heap (slot wtail - 1, ptail - 1) in registers:
    // a pop starts from it without reading memory.  Every push and pop keeps
    // it current.
    double wLk, pLk; int wLi, pLi;

    __device__ __forceinline__ void hset(int i, const BHeapEnt& e)
    {
        if (i < kHeapLds) { lk[(i - 1) * lstride] = e.key; li[(i - 1) * lstride] = (uint16_t)e.idx; }
        else wh[bslot(i)] = e;
    }
    // The global-level ancestors (index >= kHeapLds) of the two slots a
    // propagation's pushes fill, A = wtail and A + 1, read at the start of
    // the propagation together with its face and table reads, so that the
    // sift-ups compare against registers instead of waiting on memory level
    // by level.  chA[s] holds the entry at A >> (s + 1), chB[s] the one at
    // (A + 1) >> (s + 1) where that differs from A's ancestor (where they
    // coincide, chA[s] serves both).  A push into A keeps chA current, so
    // the push into A + 1 after it still reads current entries; a window
    // pop, or the push into A + 1, ends the hint (hintA = -1).
    static constexpr int kChain = 4;          // ancestors kept: every global ancestor of a slot up to 1023
    int hintA = -1;
    // chA<s> / chB<s> as named members: constant indices keep them in registers
    BHeapEnt cA0, cA1, cA2, cA3, cB0, cB1, cB2, cB3;
    template <int S> __device__ __forceinline__ BHeapEnt& chA()
    {
        if constexpr (S == 0) return cA0;
        else if constexpr (S == 1) return cA1;
        else if constexpr (S == 2) return cA2;
        else return cA3;
    }
    template <int S> __device__ __forceinline__ BHeapEnt& chB()
    {
        if constexpr (S == 0) return cB0;
        else if constexpr (S == 1) return cB1;
        else if constexpr (S == 2) return cB2;
        else return cB3;
    }
    // After the push into A, the push into A + 1 reads chA where the two
    // chains share an ancestor, so the push into A keeps chA current: the
    // copy of ancestor S follows what is written to its slot.
    template <int S> __device__ __forceinline__ void chain_set1(int pos, const BHeapEnt& e)
    {
        BHeapEnt& x = chA<S>();
        const bool m = pos == (hintA >> (S + 1));
        x.key = m ? e.key : x.key;
        x.idx = m ? e.idx : x.idx;
    }
    __device__ __forceinline__ void chain_set(int pos, const BHeapEnt& e)
    {
        chain_set1<0>(pos, e); chain_set1<1>(pos, e); chain_set1<2>(pos, e);
        chain_set1<3>(pos, e);
    }
    template <int S> __device__ __forceinline__ void hint1(int A)
    {
        const int pa = A >> (S + 1), pb = (A + 1) >> (S + 1);
        if (pa >= kHeapLds) chA<S>() = wh[bslot(pa)];
        if (pb >= kHeapLds && pb != pa) chB<S>() = wh[bslot(pb)];
    }
    // the parent of the next pseudo-heap slot when it lies in the global
    // part, read with the same batch
    int pHintI = -1;
    double pHk; int32_t pHi;
    __device__ __forceinline__ void push_hint()
    {
        const int A = wtail;
        hintA = A;
        pHintI = ptail / 2;
        if (pHintI >= kHeapLds) {
            const uint4 v = *reinterpret_cast<const uint4*>(pgh() + bslot(pHintI));
            pHk = __hiloint2double((int)v.y, (int)v.x);
            pHi = (int32_t)v.z;
        }
        hint1<0>(A); hint1<1>(A); hint1<2>(A); hint1<3>(A);
    }
    // One step of a hinted sift-up: compare with ancestor slot >> (S + 1).
    template <int S> __device__ __forceinline__ void chain_step(int slot, bool useA, double key, int& cur, bool& done)
    {
        const int pp = slot >> (S + 1);
        if (!done && pp >= kHeapLds) {
            const bool inA = useA || pp == (hintA >> (S + 1));
            const BHeapEnt ea = chA<S>(), eb = chB<S>();
            BHeapEnt pe;
            pe.key = inA ? ea.key : eb.key;
            pe.idx = inA ? ea.idx : eb.idx;
            pe.pad = 0;
            if (!(pe.key > key)) {
                done = true;
            } else {
                hset(cur, pe);
                if (cur == slot) { wLk = pe.key; wLi = pe.idx; }
                // cur is the slot itself (S = 0) or ancestor S - 1
                if constexpr (S > 0) {
                    if (useA) {
                        BHeapEnt& x = chA<S - 1>();
                        x.key = pe.key;
                        x.idx = pe.idx;
                    }
                }
                cur = pp;
            }
        }
    }
    // pseudo heap: the window heap's layout (logical entries 1..15 in LDS,
    // three global levels per 128-B block, 16-B (key, record) entries) over
    // a pool of 16-B records with its own free stack
    // (its storage follows the window heap's in each lane's arrays, at
    // offsets that are the same for every lane)
    int phOff, prOff, pfOff, plOff;
    __device__ __forceinline__ BHeapEnt* pgh() const { return wh + phOff; }
    __device__ __forceinline__ BPsRec* prr() const { return reinterpret_cast<BPsRec*>(wr + prOff); }
    __device__ __forceinline__ uint16_t* pfr() const { return wfree + pfOff; }
    __device__ __forceinline__ double* plk() const { return lk + plOff; }
    __device__ __forceinline__ uint16_t* pli() const { return li + plOff; }
    int ptail, pftop, pfresh, pcache, pcache2;
    double pRootKey; int pRootIdx;
    BPsRec pnrec;                    // record of the pseudo root
    uint32_t flags;
    bool dead;
    // The ten event counters of the propagation live in LDS (kLdsCounters
    // words per lane, lane-contiguous with an odd stride: no bank
    // conflicts), added to without waiting; the backtrace hop and error
    // counts stay in registers (the extraction kernel has no LDS counters).
    int* cl;
    int c_hops, c_err;

    __device__ __forceinline__ int twin(int e) const { return F[e / 3].tw[e % 3]; }
    __device__ __forceinline__ static int next(int e) { return (e % 3 == 2) ? e - 2 : e + 1; }
    __device__ __forceinline__ static int prev(int e) { return (e % 3 == 0) ? e + 2 : e - 1; }
    __device__ __forceinline__ int v0(int e) const { return F[e / 3].v[e % 3]; }
    __device__ __forceinline__ int v1(int e) const { return F[e / 3].v[(e % 3 == 2) ? 0 : e % 3 + 1]; }
    __device__ __forceinline__ double len(int e) const { return F[e / 3].l[e % 3]; }
    __device__ __forceinline__ int first_edge(int v) const { return firstE[v]; }
    __device__ __forceinline__ uint8_t vflag(int v) const { return vflags[v]; }
    __device__ __forceinline__ double cornerA(int e) const { return cA[e]; }
    __device__ __forceinline__ double cornerB(int e) const { return cB[e]; }
    __device__ __forceinline__ int nv() const { return nv_; }
    __device__ __forceinline__ int nf() const { return nf_; }
    __device__ __forceinline__ bool stop() const { return dead; }
    // The record is read as three 16-B vectors and rotated with selects, so
    // nothing is indexed dynamically in registers (no scratch).
    __device__ __forceinline__ void face3(int e0, Tri& t) const
    {
        const int f = e0 / 3, j = e0 - 3 * f;
        const uint4* p = reinterpret_cast<const uint4*>(F + f);
        const uint4 a = p[0], b = p[1], c = p[2];
        const double L0 = __hiloint2double((int)a.y, (int)a.x);
        const double L1 = __hiloint2double((int)a.w, (int)a.z);
        const double L2 = __hiloint2double((int)b.y, (int)b.x);
        const int V0 = (int)b.z, V1 = (int)b.w, V2 = (int)c.x;
        const int T0 = (int)c.y, T1 = (int)c.z, T2 = (int)c.w;
        const bool j0 = j == 0, j1 = j == 1;
        t.e0 = e0;
        t.e1 = 3 * f + (j0 ? 1 : (j1 ? 2 : 0));
        t.e2 = 3 * f + (j0 ? 2 : (j1 ? 0 : 1));
        t.l0 = j0 ? L0 : (j1 ? L1 : L2);
        t.l1 = j0 ? L1 : (j1 ? L2 : L0);
        t.l2 = j0 ? L2 : (j1 ? L0 : L1);
        t.a = j0 ? V0 : (j1 ? V1 : V2);
        t.b = j0 ? V1 : (j1 ? V2 : V0);
        t.op = j0 ? V2 : (j1 ? V0 : V1);
        t.tw1 = j0 ? T1 : (j1 ? T2 : T0);
        t.tw2 = j0 ? T2 : (j1 ? T0 : T1);
    }
    __device__ __forceinline__ void count(int which, long long n)
    {
        const int k = (int)n;
        switch (which) {
            case CNT_WIN_PUSH: atomicAdd(cl + 0, k); break;
            case CNT_WIN_POP: atomicAdd(cl + 1, k); break;
            case CNT_PSEUDO_PUSH: atomicAdd(cl + 2, k); break;
            case CNT_PSEUDO_POP: atomicAdd(cl + 3, k); break;
            case CNT_PROPAGATE: atomicAdd(cl + 4, k); break;
            case CNT_CHILD_BUILT: atomicAdd(cl + 5, k); break;
            case CNT_SPLIT_WRITE: atomicAdd(cl + 6, k); break;
            case CNT_VERT_WRITE: atomicAdd(cl + 7, k); break;
            case CNT_FAN_HALFEDGE: atomicAdd(cl + 8, k); break;
            case CNT_PSEUDO_NEIGH: atomicAdd(cl + 9, k); break;
            case CNT_BT_HOPS: c_hops += k; break;
            case CNT_ERRORS: c_err += k; flags |= SF_ERROR; break;
            default: break;
        }
    }

    // --- vertex table
    // The slot a lookup of face vertex `op` ended on (its own, or the empty
    // slot where it would be inserted), kept until the table next changes:
    // propagate writes that vertex right after the lookup without probing
    // again.  Only vput changes the table, and every vput clears it.
    int cacheV = -1; unsigned cacheH = 0; bool cacheIns = false;
    int cacheE = -1; unsigned cacheSH = 0; bool cacheSIns = false;
    __device__ __forceinline__ VInfo vresolve(BSlotV s, unsigned& h, int v, bool& found) const
    {
        found = false;
        for (;;) {
            if (s.key == v) {
                found = true;
                VInfo r;
                const uint32_t e24 = s.eb & kEnterNone;
                r.enter = e24 == kEnterNone ? -1 : (int)e24;
                r.birth = (int8_t)(s.eb >> 24);
                r.dist = s.dist;
                r.isSource = (uint8_t)(v == src);
                return r;
            }
            if (s.key == -1) return vinfo_default();
            h = (h + 1) & vmask;
            s = vt[h];
        }
    }
    __device__ __forceinline__ VInfo vget(int v) const
    {
        unsigned h = vslot((unsigned)v) & vmask;
        bool f;
        return vresolve(vt[h], h, v, f);
    }
    // The source keeps birth time 0 from Initialize on: no relaxation can
    // improve a distance of 0.
    // The births of the last two vertices asked for are kept in registers
    // (the next event's checks ask for the pseudo top's vertex and the
    // window top's pseudo-source, which change rarely); vput keeps them
    // current, the table reset clears them.
    int bcV0 = -1, bcB0 = 0, bcV1 = -1, bcB1 = 0;
    __device__ __forceinline__ int birth_of(int v)
    {
        if (v == src) return 0;
        if (v == bcV0) return bcB0;
        if (v == bcV1) return bcB1;
        const int b = (int)vget(v).birth;
        bcV1 = bcV0; bcB1 = bcB0;
        bcV0 = v; bcB0 = b;
        return b;
    }
    // A probe reads the whole 32-B slot at once (key and record together).
    __device__ __forceinline__ SInfo sresolve(unsigned& h, int e, bool& found) const
    {
        found = false;
        for (;;) {
            const BSlotS x = st[h];
            if (x.key == e) {
                found = true;
                SInfo r; r.dist = x.dist; r.x = x.x; r.ps = x.ps;
                return r;
            }
            if (x.key == -1) return sinfo_default();
            h = (h + 1) & smask;
        }
    }
    // The table lookups of a propagation (its far face's three vertices and
    // the split record of the edge it enters) take their first probe before
    // the window is popped: the pop never touches the tables, so the probes
    // travel with the pop's own reads.  lookup_issue reads the four slots
    // (the face must be known), lookup_finish resolves them after the pop,
    // continuing a probe past a slot held by another key.
    struct LookPre {
        int ka, kb, ko; uint32_t ea, eb, eo; double da, db, dd;
        int ks, ps; double sd, sx;
        unsigned ha, hb, ho, hs;
    };
    __device__ __forceinline__ void ld_slot(unsigned h, int& k, uint32_t& e, double& d) const
    {
        const uint4 v = *reinterpret_cast<const uint4*>(vt + h);
        k = (int)v.x; e = v.y; d = __hiloint2double((int)v.w, (int)v.z);
    }
    __device__ __forceinline__ void lookup_issue(const Tri& t, LookPre& q) const
    {
        q.ha = vslot((unsigned)t.a) & vmask;
        q.hb = vslot((unsigned)t.b) & vmask;
        q.ho = vslot((unsigned)t.op) & vmask;
        q.hs = hmix((unsigned)t.e0) & smask;
        ld_slot(q.ha, q.ka, q.ea, q.da); ld_slot(q.hb, q.kb, q.eb, q.db); ld_slot(q.ho, q.ko, q.eo, q.dd);
        {
            const uint4 v = *reinterpret_cast<const uint4*>(st + q.hs);
            const double2 w = *reinterpret_cast<const double2*>(reinterpret_cast<const char*>(st + q.hs) + 16);
            q.ks = (int)v.x; q.ps = (int)v.y; q.sd = __hiloint2double((int)v.w, (int)v.z); q.sx = w.x;
        }
    }
    __device__ __forceinline__ SInfo sresolve_from(BSlotS x, unsigned& h, int e, bool& found) const
    {
        found = false;
        for (;;) {
            if (x.key == e) {
                found = true;
                SInfo r; r.dist = x.dist; r.x = x.x; r.ps = x.ps;
                return r;
            }
            if (x.key == -1) return sinfo_default();
            h = (h + 1) & smask;
            x = st[h];
        }
    }
    __device__ __forceinline__ void lookup_finish(const Tri& t, LookPre& q, VInfo& va, VInfo& vb, VInfo& vo, SInfo& sp)
    {
        bool fa, fb, fo, fs;
        unsigned ha = q.ha, hb = q.hb, ho = q.ho, hs = q.hs;
        BSlotS ss; ss.key = q.ks; ss.ps = q.ps; ss.dist = q.sd; ss.x = q.sx; ss.pad = 0.0;
        sp = sresolve_from(ss, hs, t.e0, fs);
        BSlotV sa, sb, so;
        sa.key = q.ka; sa.eb = q.ea; sa.dist = q.da;
        sb.key = q.kb; sb.eb = q.eb; sb.dist = q.db;
        so.key = q.ko; so.eb = q.eo; so.dist = q.dd;
        va = vresolve(sa, ha, t.a, fa);
        vb = vresolve(sb, hb, t.b, fb);
        vo = vresolve(so, ho, t.op, fo);
        cacheV = t.op; cacheH = ho; cacheIns = !fo;
        cacheE = t.e0; cacheSH = hs; cacheSIns = !fs;
    }
    // --- fans (FanRec): a pseudo-source's sweeps, windows and neighbours
    __device__ __forceinline__ bool fan_ok() const { return fanR != nullptr; }
    __device__ __forceinline__ void fan_base(int v, int& base, int& d) const
    {
        base = fanBase[v];
        d = fanBase[v + 1] - base;
    }
    __device__ __forceinline__ int fan_index(int e) const { return fanIdx[e]; }
    __device__ __forceinline__ double fan_cA(int j) const { return fanR[j].cA; }
    __device__ __forceinline__ double fan_cB(int j) const { return fanR[j].cB; }
    __device__ __forceinline__ void fan_window(int j, double& lo, double& lnx, double& lpv, int& o, int& twn) const
    {
        const double2 x = *reinterpret_cast<const double2*>(&fanR[j].lo);
        const uint4 y = *reinterpret_cast<const uint4*>(&fanR[j].lpv);
        lo = x.x; lnx = x.y;
        lpv = __hiloint2double((int)y.y, (int)y.x);
        o = (int)y.z; twn = (int)y.w;
    }
    __device__ __forceinline__ void fan_neigh(int j, int& nb, int& tw, int& flag, double& lo) const
    {
        const uint4 y = *reinterpret_cast<const uint4*>(&fanR[j].nb);
        nb = (int)y.x; tw = (int)y.y; flag = (int)y.z;
        lo = fanR[j].lo;
    }
    // a vertex lookup in two halves: the first probe's read, then its resolution
    struct VProbe { unsigned h; int k; uint32_t e; double d; };
    __device__ __forceinline__ void vprobe_issue(int v, VProbe& q) const
    {
        q.h = vslot((unsigned)v) & vmask;
        ld_slot(q.h, q.k, q.e, q.d);
    }
    __device__ __forceinline__ VInfo vprobe_finish(int v, const VProbe& q) const
    {
        BSlotV s; s.key = q.k; s.eb = q.e; s.dist = q.d;
        unsigned h = q.h;
        bool f;
        return vresolve(s, h, v, f);
    }
    __device__ __forceinline__ void vput(int v, const VInfo& x)
    {
        if (dead) return;
        unsigned h;
        bool ins;
        if (v == cacheV) {
            h = cacheH;
            ins = cacheIns;
        } else {
            h = vslot((unsigned)v) & vmask;
            for (;;) {
                const int32_t k = vt[h].key;
                if (k == v) { ins = false; break; }
                if (k == -1) { ins = true; break; }
                h = (h + 1) & vmask;
            }
        }
        cacheV = -1;
        if (ins) {
            if (vcount >= vlimit) { flags |= SF_CAPACITY; dead = true; return; }
            rl[vcount++] = v;
            if (vcount > vrefcap) flags |= SF_REF_OVERFLOW;
        }
        BSlotV ns;
        ns.key = v;
        ns.eb = ((uint32_t)x.enter & kEnterNone) | ((uint32_t)(uint8_t)x.birth << 24);
        ns.dist = x.dist;
        vt[h] = ns;
        if (v == bcV0) bcB0 = (int)x.birth;
        if (v == bcV1) bcB1 = (int)x.birth;
    }
    __device__ __forceinline__ void vtouch(int) {}
    // --- split table
    __device__ __forceinline__ SInfo sget(int e) const
    {
        unsigned h = hmix((unsigned)e) & smask;
        bool f;
        return sresolve(h, e, f);
    }
    __device__ __forceinline__ void sput(int e, const SInfo& x)
    {
        if (dead) return;
        unsigned h;
        bool ins;
        if (e == cacheE) {
            h = cacheSH;
            ins = cacheSIns;
        } else {
            h = hmix((unsigned)e) & smask;
            for (;;) {
                const int32_t k = st[h].key;
                if (k == e) { ins = false; break; }
                if (k == -1) { ins = true; break; }
                h = (h + 1) & smask;
            }
        }
        cacheE = -1;
        if (ins) {
            if (scount >= slimit) { flags |= SF_CAPACITY; dead = true; return; }
            ++scount;
            if (scount > srefcap) flags |= SF_REF_OVERFLOW;
        }
        BSlotS ns;
        ns.key = e; ns.ps = x.ps; ns.dist = x.dist; ns.x = x.x; ns.pad = 0.0;
        st[h] = ns;
    }

    // --- window heap: PriorityQueue.cuh:51-105 semantics on (key, record) entries
    // Free record slots form a stack whose top two entries are kept in
    // registers (wcache, then wcache2, then wfree[0..wftop)); the refill
    // read after an allocation is needed only two allocations later, so
    // the two pushes of one propagation never wait on it.  A slot below
    // the register part exists only while both registers are filled.
    __device__ __forceinline__ int walloc()
    {
        int idx;
        if (wcache >= 0) {
            idx = wcache;
            wcache = wcache2;
            wcache2 = wftop > 0 ? (int)wfree[--wftop] : -1;
        } else {
            idx = wfresh++;
        }
        return idx;
    }
    __device__ __forceinline__ void wrelease(int idx)
    {
        if (wcache2 >= 0) wfree[wftop++] = (uint16_t)wcache2;
        wcache2 = wcache;
        wcache = idx;
    }
    __device__ __forceinline__ bool wempty() const { return wtail == 1; }
    __device__ __forceinline__ void wtop(Win& w) const
    {
        // Every change of the root refreshes nrec (pop_either loads it, wpush
        // copies the new record), so the root's record is always at hand.
        const BWinRec& r = nrec;
        w.b0 = r.b0; w.b1 = r.b1; w.ps = r.ps; w.sx = r.sx; w.sy = r.sy;
        w.d0 = 0.0; w.d1 = 0.0;
        w.minDist = rootKey;
        w.edge = r.edge; w.pseudoSrc = r.pseudoSrc; w.twin = r.twin;
        w.birth = (int)(int8_t)(r.lvlb & 0xFF);
        w.level = r.lvlb >> 8;
    }
    // LDS-only and global-only reads of heap entries: with one accessor
    // for both, the compiler merges the two paths into generic (flat)
    // loads, whose waits also wait for every global load in flight.
    __device__ __forceinline__ double lkey(int i) const { return lk[(i - 1) * lstride]; }
    __device__ __forceinline__ int lidx(int i) const { return li[(i - 1) * lstride]; }
    __device__ __forceinline__ BHeapEnt gget(int i) const { return wh[bslot(i)]; }
    __device__ __forceinline__ void wpush(const Win& w)
    {
        if (wtail - 1 >= wcap) {
            // at the reference capacity a push is dropped (PriorityQueue.cuh:53); below it the
            // source outgrew its storage and is recomputed on the plain path
            if (wcap >= kHeapCap) flags |= SF_WIN_DROP;
            else { flags |= SF_CAPACITY; dead = true; }
            return;
        }
        const int idx = walloc();
        BWinRec rec;
        rec.b0 = w.b0; rec.b1 = w.b1; rec.ps = w.ps; rec.sx = w.sx; rec.sy = w.sy;
        rec.edge = w.edge; rec.pseudoSrc = w.pseudoSrc;
        rec.lvlb = (w.level << 8) | (w.birth & 0xFF); rec.twin = w.twin;
        wr[idx] = rec;
        ++wtail;
        const int slot = wtail - 1;
        wmax = slot > wmax ? slot : wmax;
        int cur = slot, p = cur / 2;
        const double key = w.minDist;
        BHeapEnt ne; ne.key = key; ne.idx = idx; ne.pad = 0;
        // global levels of the path (parents with index >= 16), then LDS ones
        bool done = false;
        const int A = hintA;
        const bool useA = A >= 0 && slot == A;
        const bool useB = A >= 0 && slot == A + 1;
        if (useA || useB) {
            // step S compares with the ancestor slot >> (S + 1) from its
            // register copy (hset keeps the copies current as entries move)
            chain_step<0>(slot, useA, key, cur, done);
            chain_step<1>(slot, useA, key, cur, done);
            chain_step<2>(slot, useA, key, cur, done);
            chain_step<3>(slot, useA, key, cur, done);
            p = cur / 2;
            // ancestors above the kept ones (none below slot 1024)
            while (!done && p >= kHeapLds) {
                const BHeapEnt pe = gget(p);
                if (!(pe.key > key)) { done = true; break; }
                if (useA) chain_set(cur, pe);
                hset(cur, pe);
                if (cur == slot) { wLk = pe.key; wLi = pe.idx; }
                cur = p; p = cur / 2;
            }
        } else {
            while (p >= kHeapLds) {
                const BHeapEnt pe = gget(p);
                if (!(pe.key > key)) { done = true; break; }
                hset(cur, pe);
                if (cur == slot) { wLk = pe.key; wLi = pe.idx; }
                cur = p; p = cur / 2;
            }
        }
        if (!done) {
            while (p > 0) {
                const double pk = lkey(p);
                if (!(pk > key)) break;
                BHeapEnt pe; pe.key = pk; pe.idx = lidx(p); pe.pad = 0;
                if (useA && cur >= kHeapLds) chain_set(cur, pe);
                hset(cur, pe);
                if (cur == slot) { wLk = pe.key; wLi = pe.idx; }
                cur = p; p = cur / 2;
            }
        }
        if (useA && cur >= kHeapLds) chain_set(cur, ne);
        hset(cur, ne);
        if (cur == slot) { wLk = key; wLi = idx; }
        if (useB) hintA = -1;
        if (cur == 1) { rootKey = key; rootIdx = idx; nrec = rec; nrecIdx = idx; }
    }
    // --- one pop for either queue.  Both queues share a layout, so lanes
    // that pop their window queue and lanes that pop their pseudo queue run
    // the same instructions at once; `isP` picks the queue's storage (its
    // offsets), tail, root copies, record and free stack.
    __device__ __forceinline__ void load_block_at(int base, double (&k)[8], int (&ix)[8]) const
    {
        const uint4* q = reinterpret_cast<const uint4*>(wh + base);
#pragma unroll
        for (int s = 1; s < 8; ++s) {
            const uint4 v = q[s];
            k[s] = __hiloint2double((int)v.y, (int)v.x);
            ix[s] = (int)v.z;
        }
    }
    __device__ __forceinline__ void hset_at(int i, const BHeapEnt& e, int goff, int loff)
    {
        if (i < kHeapLds) { lk[loff + (i - 1) * lstride] = e.key; li[loff + (i - 1) * lstride] = (uint16_t)e.idx; }
        else wh[goff + bslot(i)] = e;
    }
    __device__ __forceinline__ bool sift_blocks_at(int& cur, const BHeapEnt& top, int blkOfLeft, int goff, int loff,
                                                   int tail)
    {
        double ka[8], kb[8];
        int ia[8], ib[8];
        load_block_at(goff + blkOfLeft * 8, ka, ia);
        load_block_at(goff + (blkOfLeft + 1) * 8, kb, ib);
        int l = 2 * cur, r = l + 1;
        if (l >= tail) return false;
        bool right = r < tail && !(ka[1] < kb[1]);
        double ck = right ? kb[1] : ka[1];
        int ci = right ? ib[1] : ia[1];
        if (!(ck < top.key)) return false;
        { BHeapEnt m; m.key = ck; m.idx = ci; m.pad = 0; hset_at(cur, m, goff, loff); }
        cur = right ? r : l;
        double x[8];
        int xi[8];
#pragma unroll
        for (int s = 2; s < 8; ++s) { x[s] = right ? kb[s] : ka[s]; xi[s] = right ? ib[s] : ia[s]; }
        l = 2 * cur; r = l + 1;
        if (l >= tail) return false;
        right = r < tail && !(x[2] < x[3]);
        ck = right ? x[3] : x[2];
        ci = right ? xi[3] : xi[2];
        if (!(ck < top.key)) return false;
        { BHeapEnt m; m.key = ck; m.idx = ci; m.pad = 0; hset_at(cur, m, goff, loff); }
        cur = right ? r : l;
        const double y0 = right ? x[6] : x[4], y1 = right ? x[7] : x[5];
        const int j0 = right ? xi[6] : xi[4], j1 = right ? xi[7] : xi[5];
        l = 2 * cur; r = l + 1;
        if (l >= tail) return false;
        right = r < tail && !(y0 < y1);
        ck = right ? y1 : y0;
        ci = right ? j1 : j0;
        if (!(ck < top.key)) return false;
        { BHeapEnt m; m.key = ck; m.idx = ci; m.pad = 0; hset_at(cur, m, goff, loff); }
        cur = right ? r : l;
        return true;
    }
    // Sift-down through levels 8-9: the children of `cur` (level 7) are the
    // level-8 roots of two adjacent 64-B blocks, each with its two children.
    __device__ __forceinline__ void sift_deep_at(int& cur, const BHeapEnt& top, int dblkOfLeft, int goff, int loff,
                                                 int tail)
    {
        double ka[3], kb[3];
        int ia[3], ib[3];
        const uint4* qa = reinterpret_cast<const uint4*>(wh + goff + kDeepBase + dblkOfLeft * 4);
        const uint4* qb = qa + 4;
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            const uint4 va = qa[s], vb = qb[s];
            ka[s] = __hiloint2double((int)va.y, (int)va.x); ia[s] = (int)va.z;
            kb[s] = __hiloint2double((int)vb.y, (int)vb.x); ib[s] = (int)vb.z;
        }
        int l = 2 * cur, r = l + 1;
        if (l >= tail) return;
        bool right = r < tail && !(ka[0] < kb[0]);
        double ck = right ? kb[0] : ka[0];
        int ci = right ? ib[0] : ia[0];
        if (!(ck < top.key)) return;
        { BHeapEnt m; m.key = ck; m.idx = ci; m.pad = 0; hset_at(cur, m, goff, loff); }
        cur = right ? r : l;
        const double y0 = right ? kb[1] : ka[1], y1 = right ? kb[2] : ka[2];
        const int j0 = right ? ib[1] : ia[1], j1 = right ? ib[2] : ia[2];
        l = 2 * cur; r = l + 1;
        if (l >= tail) return;
        right = r < tail && !(y0 < y1);
        ck = right ? y1 : y0;
        ci = right ? j1 : j0;
        if (!(ck < top.key)) return;
        { BHeapEnt m; m.key = ck; m.idx = ci; m.pad = 0; hset_at(cur, m, goff, loff); }
        cur = right ? r : l;
    }
    // PriorityQueue.cuh:75-105 on the window queue (isP false) or the
    // pseudo queue (isP true)
    __device__ __forceinline__ void pop_either(bool isP)
    {
        if (isP) { pHintI = -1; prelease(pRootIdx); }
        else { hintA = -1; wrelease(rootIdx); }
        const int goff = isP ? phOff : 0;
        const int loff = isP ? plOff : 0;
        const int tail = (isP ? ptail : wtail) - 1;
        if (isP) ptail = tail; else wtail = tail;
        const double tk = isP ? pLk : wLk;
        const int ti = isP ? pLi : wLi;
        // The popped heap's last entry (slot `tail`) comes from its register
        // copy.  The entry at slot tail - 1 becomes the last one: it has no
        // children below the new tail, so the sift-down either ends there
        // (and leaves `top` in it) or does not touch it.  Its value is read
        // now, with the pop's other reads, and chosen after the sift.
        const int nl = tail - 1;
        double nlk;
        int nli;
        {
            const int il = nl < kHeapLds ? (nl > 0 ? nl : 1) : kHeapLds - 1;
            const BHeapEnt g = wh[goff + bslot(nl < kHeapLds ? kHeapLds : nl)];
            const double lkk = lk[loff + (il - 1) * lstride];
            const int lii = li[loff + (il - 1) * lstride];
            nlk = nl < kHeapLds ? lkk : g.key;
            nli = nl < kHeapLds ? lii : g.idx;
        }
        BHeapEnt top; top.key = tk; top.idx = ti; top.pad = 0;
        int cur = 1;
        int mc = -1;
        double mk = 0.0;
        int mi = 0;
        {
            const int l = 2, r = 3;
            if (r < tail) {
                const double kl = lk[loff + (l - 1) * lstride], kr = lk[loff + (r - 1) * lstride];
                const int il2 = li[loff + (l - 1) * lstride], ir2 = li[loff + (r - 1) * lstride];
                const bool left = kl < kr;
                mc = left ? l : r; mk = left ? kl : kr; mi = left ? il2 : ir2;
            } else if (l < tail) {
                mc = l; mk = lk[loff + (l - 1) * lstride]; mi = li[loff + (l - 1) * lstride];
            }
        }
        // the new root is known from the first level: its record is loaded
        // now and arrives while the sift continues
        const bool rootMoves = mc != -1 && mk < top.key;
        const double nk = rootMoves ? mk : top.key;
        const int ni = rootMoves ? mi : top.idx;
        if (isP) {
            pRootKey = nk; pRootIdx = ni;
            if (tail > 1) pnrec = prr()[ni];
        } else {
            rootKey = nk; rootIdx = ni;
            if (tail > 1) { nrec = wr[ni]; nrecIdx = ni; }
            else nrecIdx = -1;
        }
        bool deeper = false;
        while (mc != -1 && mk < top.key) {
            lk[loff + (cur - 1) * lstride] = mk;
            li[loff + (cur - 1) * lstride] = (uint16_t)mi;
            cur = mc;
            if (cur >= 16) { deeper = true; break; }
            const int l = 2 * cur, r = l + 1;
            if (r < tail) {
                const double kl = lk[loff + (l - 1) * lstride], kr = lk[loff + (r - 1) * lstride];
                const int il2 = li[loff + (l - 1) * lstride], ir2 = li[loff + (r - 1) * lstride];
                const bool left = kl < kr;
                mc = left ? l : r; mk = left ? kl : kr; mi = left ? il2 : ir2;
            } else if (l < tail) {
                mc = l; mk = lk[loff + (l - 1) * lstride]; mi = li[loff + (l - 1) * lstride];
            } else {
                mc = -1;
            }
        }
        if (deeper && sift_blocks_at(cur, top, 2 * cur - 32, goff, loff, tail) && 2 * cur < tail)
            sift_deep_at(cur, top, 2 * cur - 256, goff, loff, tail);
        hset_at(cur, top, goff, loff);
        const double nk2 = cur == nl ? top.key : nlk;
        const int ni2 = cur == nl ? top.idx : nli;
        if (isP) { pLk = nk2; pLi = ni2; } else { wLk = nk2; wLi = ni2; }
    }
    // --- pseudo heap: the window heap's operations on the pseudo storage
    __device__ __forceinline__ int palloc()
    {
        int idx;
        if (pcache >= 0) {
            idx = pcache;
            pcache = pcache2;
            pcache2 = pftop > 0 ? (int)pfr()[--pftop] : -1;
        } else {
            idx = pfresh++;
        }
        return idx;
    }
    __device__ __forceinline__ void prelease(int idx)
    {
        if (pcache2 >= 0) pfr()[pftop++] = (uint16_t)pcache2;
        pcache2 = pcache;
        pcache = idx;
    }
    __device__ __forceinline__ double plkey(int i) const { return plk()[(i - 1) * lstride]; }
    __device__ __forceinline__ int plidx(int i) const { return pli()[(i - 1) * lstride]; }
    __device__ __forceinline__ BHeapEnt pgget(int i) const { return pgh()[bslot(i)]; }
    __device__ __forceinline__ void phset(int i, const BHeapEnt& e)
    {
        if (i < kHeapLds) { plk()[(i - 1) * lstride] = e.key; pli()[(i - 1) * lstride] = (uint16_t)e.idx; }
        else pgh()[bslot(i)] = e;
    }
    __device__ __forceinline__ bool pempty() const { return ptail == 1; }
    __device__ __forceinline__ void ptop(Pseudo& q) const
    {
        q.dist = pRootKey; q.vert = pnrec.vert; q.pseudoSrc = pnrec.pseudoSrc;
        q.birth = (int)(int8_t)(pnrec.lvlb & 0xFF); q.level = pnrec.lvlb >> 8;
    }
    // PriorityQueue.cuh:51-66 on the pseudo heap (as wpush); pops go through pop_either
    __device__ __forceinline__ void ppush(const Pseudo& q)
    {
        if (ptail - 1 >= pcap) {
            if (pcap >= kHeapCap) flags |= SF_PSEUDO_DROP;
            else { flags |= SF_CAPACITY; dead = true; }
            return;
        }
        const int idx = palloc();
        BPsRec rec;
        rec.vert = q.vert; rec.pseudoSrc = q.pseudoSrc;
        rec.lvlb = (q.level << 8) | (q.birth & 0xFF); rec.pad = 0;
        prr()[idx] = rec;
        ++ptail;
        const int slot = ptail - 1;
        pmax = slot > pmax ? slot : pmax;
        int cur = slot, p = cur / 2;
        const double key = q.dist;
        BHeapEnt ne; ne.key = key; ne.idx = idx; ne.pad = 0;
        const int hinted = pHintI;
        pHintI = -1;
        bool done = false;
        // the first global level from the parent read ahead (its own block
        // of code: no choice between a register copy and a memory read)
        if (p >= kHeapLds && p == hinted) {
            if (!(pHk > key)) {
                done = true;
            } else {
                BHeapEnt pe; pe.key = pHk; pe.idx = pHi; pe.pad = 0;
                phset(cur, pe);
                pLk = pe.key; pLi = pe.idx;          // cur == slot here
                cur = p; p = cur / 2;
            }
        }
        while (!done && p >= kHeapLds) {
            const BHeapEnt pe = pgget(p);
            if (!(pe.key > key)) { done = true; break; }
            phset(cur, pe);
            if (cur == slot) { pLk = pe.key; pLi = pe.idx; }
            cur = p; p = cur / 2;
        }
        if (!done) {
            while (p > 0) {
                const double pk = plkey(p);
                if (!(pk > key)) break;
                BHeapEnt pe; pe.key = pk; pe.idx = plidx(p); pe.pad = 0;
                phset(cur, pe);
                if (cur == slot) { pLk = pe.key; pLi = pe.idx; }
                cur = p; p = cur / 2;
            }
        }
        phset(cur, ne);
        if (cur == slot) { pLk = key; pLi = idx; }
        if (cur == 1) { pRootKey = key; pRootIdx = idx; pnrec = rec; }
    }
};

struct BArgs {
    DevMesh m;
    DevGraph g;
    int K;
    int lanes;
    int nsrc;                   // sources in this chunk
    int srcBase;                // first source of the chunk
    unsigned vcap, scap;        // table slots (powers of two)
    int vrefcap, srefcap;       // reference table sizes (vertCoef*K, splitCoef*K)
    BSlotV* vtab;               // chunk * vcap
    int32_t* rlist;             // chunk * vcap
    BSlotS* stab;               // chunk * scap
    BHeapEnt* wheap;            // lanes * (hslots + phslots): window heap, then pseudo heap (blocked layout)
    BWinRec* wrec;              // lanes * (wcap + 1 + precUnits): window records, then pseudo records
    uint16_t* wfree;            // lanes * (wcap + pcap + 2): window free stack, then pseudo free stack
    int* work;                  // work counter
    uint32_t* status;           // chunk
    int32_t* vcount;            // chunk
    int32_t* cnt;               // chunk * kNumBatchedCounters event counters
    int hopCap;
    const int32_t* order;       // sources in processing order (null: by id)
    int pseudoShare;            // pseudo events run when nP * pseudoShare >= the wave's stepping lanes
    int wcap, pcap;             // window / pseudo queue storage per lane (kHeapCap: the reference's)
    int hslots;                 // window heap slots per lane (blocked layout)
    int phslots;                // pseudo heap slots per lane, after the window heap's
    int precUnits;              // 64-B units of pseudo records per lane, after the window records
    int lanesPerWave;           // source lanes per 64-wide wave
};

constexpr int kNumBatchedCounters = 15;
constexpr int kLdsCounters = 10;            // event counters 0..9 of a propagation lane, in LDS
constexpr int kLdsCounterStride = 11;       // words per lane (odd: lanes hit distinct banks)
// Launch shapes the autotuner tries, the default one first: threads per
// block, source lanes per wave.  Each keeps one block's LDS (664 B per
// lane) within 64 KiB, so none is cut down to another.
constexpr int kTuneShapes = 5;
constexpr int kTuneShape[kTuneShapes][2] = {{128, 28}, {64, 28}, {128, 30}, {128, 26}, {64, 32}};
// A wave runs its pseudo-source events once at least 1/kPseudoShare of its
// stepping lanes wait for one.
constexpr int kPseudoShare = 4;
// Source lanes per wave and threads per block when neither is set.  With
// one source per lane the chip holds about one wave per SIMD, so a wave's
// memory waits go unhidden; partly filled waves give two waves per SIMD
// that overlap each other's waits, at the price of issuing each
// instruction for fewer sources.  On the 72K torus at K=500, 28 lanes is
// the fastest (24: +16%, 26: +3%, 30: +2%, 32: +3.5%, 36: +5.5%).
constexpr int kLanesPerWave = 28;
constexpr int kPropBlock = 128;
// Without recorded event counts, the predicted order is used while the
// sources are at most this many per lane: the order decides how long the
// launch's tail is when a lane runs one or two sources, while with many
// per lane the shared work counter already evens the lanes out (no
// difference measured at 13 per lane).
constexpr int kPredictPerLane = 4;
// Weight of a pseudo-source pop against a window pop in the recorded cost
// that orders sources: the wave iterations that run fan sweeps take about
// three times as long as those that run window events only.
constexpr int kPseudoCost = 3;

__device__ __forceinline__ void policy_mesh(BatchedPolicy& p, const BArgs& a)
{
    p.F = a.m.face; p.fanR = a.m.fan; p.fanIdx = a.m.fanIdx; p.fanBase = a.m.fanBase; p.cA = a.m.cornerA; p.cB = a.m.cornerB; p.firstE = a.m.firstEdge;
    p.vflags = a.m.flags; p.nv_ = a.m.nv; p.nf_ = a.m.nf;
}

__device__ __forceinline__ void policy_tables(BatchedPolicy& p, const BArgs& a, int slot, int src)
{
    p.src = src;
    p.vt = a.vtab + (size_t)slot * a.vcap;
    p.rl = a.rlist + (size_t)slot * a.vcap;
    p.vmask = a.vcap - 1;
    p.st = a.stab + (size_t)slot * a.scap;
    p.smask = a.scap - 1;
    p.vrefcap = a.vrefcap; p.srefcap = a.srefcap;
    p.vlimit = (int)(a.vcap - (a.vcap >> 2));      // keep the load factor <= 3/4
    p.slimit = (int)(a.scap - (a.scap >> 2));
    p.vcount = 0; p.scount = 0;
    p.bcV0 = -1; p.bcV1 = -1;
    p.cacheV = -1; p.cacheE = -1;
    p.hintA = -1;
    p.pHintI = -1;
    p.flags = 0; p.dead = false;
    p.c_hops = p.c_err = 0;
}

__global__ void __launch_bounds__(256) batched_prop_kernel(BArgs a)
{
    // (kHeapLds-1) window keys, then as many pseudo keys, per source lane;
    // then the window and the pseudo indices
    extern __shared__ double ldsHeap[];
    // the first lanesPerWave lanes of each wave run sources; the LDS heap
    // levels are laid out for those lanes only
    const int wl = threadIdx.x & 63;
    if (wl >= a.lanesPerWave) return;
    const int lanesPerBlock = (int)(blockDim.x >> 6) * a.lanesPerWave;
    const int bl = (int)(threadIdx.x >> 6) * a.lanesPerWave + wl;
    const int lane = (int)blockIdx.x * lanesPerBlock + bl;
    if (lane >= a.lanes) return;
    BatchedPolicy p;
    policy_mesh(p, a);
    p.lstride = lanesPerBlock;
    p.lk = ldsHeap + bl;
    p.li = reinterpret_cast<uint16_t*>(ldsHeap + 2 * (kHeapLds - 1) * lanesPerBlock) + bl;
    p.plOff = (kHeapLds - 1) * lanesPerBlock;
    p.cl = reinterpret_cast<int*>(reinterpret_cast<char*>(ldsHeap) +
                                  (size_t)2 * (kHeapLds - 1) * lanesPerBlock * (sizeof(double) + sizeof(uint16_t))) +
           bl * kLdsCounterStride;
    p.wh = a.wheap + (size_t)lane * (a.hslots + a.phslots);
    p.wr = a.wrec + (size_t)lane * (a.wcap + 1 + a.precUnits);
    p.wfree = a.wfree + (size_t)lane * (a.wcap + a.pcap + 2);
    p.phOff = a.hslots; p.prOff = a.wcap + 1; p.pfOff = a.wcap + 1;
    p.wcap = a.wcap; p.pcap = a.pcap;
    using I = Ich<BatchedPolicy>;
    I ich(p, 0, a.K);
    Win top;
    Pseudo ptop;
    int slot = -1, ev = I::EV_DONE;
    bool have = false;
    // Each iteration every lane of the wave takes one step of its source.  A
    // pseudo-source event (a fan sweep, the longer path) waits until enough
    // of the wave's lanes want one, so the wave runs that path for many
    // lanes at once instead of in nearly every iteration; a source's own
    // order of operations is unchanged.  (A stale pseudo pop does not wait:
    // it stands before every later step of its source.)
    for (;;) {
        if (!have) {
            slot = atomicAdd(a.work, 1);
            if (slot >= a.nsrc) break;
            const int src = a.order ? a.order[a.srcBase + slot] : a.srcBase + slot;
            policy_tables(p, a, slot, src);
#pragma unroll
            for (int i = 0; i < kLdsCounters; ++i) p.cl[i] = 0;
            p.wtail = 1; p.wftop = 0; p.wfresh = 0; p.wcache = -1; p.wcache2 = -1; p.rootKey = 0.0; p.rootIdx = 0;
            p.wmax = 0; p.pmax = 0;
            p.nrecIdx = -1;
            p.ptail = 1; p.pftop = 0; p.pfresh = 0; p.pcache = -1; p.pcache2 = -1;
            ich.reset(src);
            ich.initialize();
            ev = ich.next_event(top, ptop);
            have = true;
        }
        if (ev >= I::EV_STALE_W && ev <= I::EV_PSEUDO) {
            const bool wantP = ev == I::EV_PSEUDO;
            const int nP = __popcll(__ballot(wantP));
            const int nAll = __popcll(__ballot(true));
            if (!wantP || nP * a.pseudoShare >= nAll) {
                if (ich.do_event(ev, top, ptop)) ev = ich.next_event(top, ptop);
                else ev = I::EV_DONE;
            }
        }
        if (ev == I::EV_DONE || ev == I::EV_STUCK) {
            if (ev == I::EV_STUCK) p.count(CNT_ERRORS, 1);
            a.status[slot] = p.flags;
            a.vcount[slot] = p.vcount;
            int32_t* c = a.cnt + (size_t)slot * kNumBatchedCounters;
#pragma unroll
            for (int i = 0; i < kLdsCounters; ++i) c[i] = p.cl[i];
            c[10] = 0; c[11] = p.vcount; c[12] = p.scount; c[13] = p.wmax; c[14] = p.pmax;
            have = false;
        }
    }
}

// One wave per source: sort the reached vertices by id in LDS, backtrace 64
// candidates at a time, keep the first K whose path passes no vertex.
template <int SORTN>
__global__ void __launch_bounds__(64) batched_extract_kernel(BArgs a)
{
    __shared__ int32_t buf[SORTN];
    const int lane = threadIdx.x;
    BatchedPolicy p;
    policy_mesh(p, a);
    for (int slot = blockIdx.x; slot < a.nsrc; slot += gridDim.x) {
        const int src = a.order ? a.order[a.srcBase + slot] : a.srcBase + slot;
        const uint32_t stt = a.status[slot];
        const int K = a.K;
        if (stt & (SF_CAPACITY | SF_REF_OVERFLOW)) {
            if (lane == 0) a.g.tail[src] = 0;
            continue;
        }
        const int n = a.vcount[slot];
        policy_tables(p, a, slot, src);
        int N = 64;
        while (N < n) N <<= 1;
        for (int i = lane; i < N; i += 64) buf[i] = i < n ? p.rl[i] : 0x7fffffff;
        __syncthreads();
        for (int k = 2; k <= N; k <<= 1) {
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int i = lane; i < N; i += 64) {
                    const int ixj = i ^ j;
                    if (ixj > i) {
                        const int x = buf[i], y = buf[ixj];
                        const bool asc = (i & k) == 0;
                        if ((x > y) == asc) { buf[i] = y; buf[ixj] = x; }
                    }
                }
                __syncthreads();
            }
        }
        Ich<BatchedPolicy> ich(p, src, K);
        int tail = 0;
        const size_t row = (size_t)src * (size_t)K;
        for (int base = 0; base < n && tail < K; base += 64) {
            const int c = base + lane;
            const int j = c < n ? buf[c] : -1;
            bool q = false;
            int se = -1, de = -1;
            double sx = 0.0, dx = 0.0, dist = 0.0;
            if (j >= 0 && j != src) {
                dist = p.vget(j).dist;
                const bool pass = ich.path_to(j, se, sx, de, dx, a.hopCap);
                q = !pass;
                if (se == -1) { se = de; sx = dx; }
            }
            const unsigned long long m = __ballot(q);
            const int pos = tail + __popcll(m & ((1ull << lane) - 1ull));
            if (q && pos < K) {
                a.g.adj[row + pos] = j;
                a.g.w[row + pos] = dist;
                a.g.srcEdge[row + pos] = se; a.g.srcX[row + pos] = sx;
                a.g.dstEdge[row + pos] = de; a.g.dstX[row + pos] = dx;
            }
            tail += __popcll(m);
        }
        const int hops = p.c_hops;
        // wave reduction of the backtrace hop count
        int hsum = hops;
        for (int off = 32; off > 0; off >>= 1) hsum += __shfl_xor(hsum, off);
        if (lane == 0) {
            a.g.tail[src] = tail < K ? tail : K;
            a.cnt[(size_t)slot * kNumBatchedCounters + 10] = hsum;
        }
        if (p.c_err) atomicOr(&a.status[slot], SF_ERROR);
        __syncthreads();
    }
}

static unsigned next_pow2(unsigned x)
{
    unsigned p = 1;
    while (p < x) p <<= 1;
    return p;
}

// LDS of one propagation block: heap levels 1..31 of both queues (8-B key
// and 16-bit record index each) and the event counters, for every source
// lane of the block.
static size_t lds_bytes(int block, int lanesPerWave)
{
    const size_t lanes = (size_t)(block / 64) * (size_t)lanesPerWave;
    return (size_t)2 * (kHeapLds - 1) * lanes * (sizeof(double) + sizeof(uint16_t)) + lanes * kLdsCounterStride * sizeof(int);
}

// Source lanes the device holds at once for a launch shape: resident
// blocks per CU (registers, LDS) x CUs x lanes per block.  More lanes than
// this would only start as resident waves end, with the work already taken.
static int resident_lanes(int block, int lanesPerWave)
{
    int perCU = 0, dev = 0, nCU = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, batched_prop_kernel, block,
                                                     lds_bytes(block, lanesPerWave)) != hipSuccess ||
        hipDeviceGetAttribute(&nCU, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || perCU <= 0 ||
        nCU <= 0)
        return 65536;
    return perCU * nCU * (block / 64) * lanesPerWave;
}

std::string build_batched(const HostMesh& h, const DevMesh& d, const BuildConfig& c, DevGraph& g,
                          BuildStats& st)
{
    const int K = c.K;
    const int V = h.nv;
    if (h.ne >= (int)kEnterNone) {
        // enter edges are packed into 24 bits of a vertex-table slot: larger
        // meshes run every source on the plain path (same rows)
        if (c.ensurePlainMesh) {
            const std::string ue = c.ensurePlainMesh();
            if (!ue.empty()) return ue;
        }
        std::string e = build_plain(h, d, c, g, st, nullptr, 0);
        st.fallbackPlain = 1;
        return e;
    }
    int lanesPerWave = c.lanesPerWave > 0 ? c.lanesPerWave : kLanesPerWave;
    if (lanesPerWave > 64) return "lanes_per_wave must be at most 64";
    if (c.queueCap < 0 || c.queueCap > kHeapCap) return "queue_cap must be in [0, 1023]";
    const int vrefcap = (int)((unsigned)K * c.vertCoef);
    const int srefcap = (int)((unsigned)K * c.splitCoef);
    unsigned vcap = c.vertCap > 0 ? next_pow2((unsigned)c.vertCap) : next_pow2(2u * (unsigned)vrefcap + 2u);
    unsigned scap = c.splitCap > 0 ? next_pow2((unsigned)c.splitCap) : next_pow2(4u * (unsigned)K);
    int block = c.propBlock > 0 ? c.propBlock : kPropBlock;
    // Queue storage per lane: the reference's 1023 items unless set; a
    // source that outgrows a smaller storage is recomputed exactly on the
    // plain path (the rows never depend on it).
    int wcap = c.queueCap > 0 ? c.queueCap : kHeapCap;
    int pcap = c.queueCap > 0 ? c.queueCap : kHeapCap;
    // Autotuner (per mesh, K and coefficients; results never depend on it):
    // once a build has recorded every source's event count, sources start
    // from the costliest down; when the default tables would
    // not fit at once they are sized to the peak occupancy recorded (a source
    // that still overflows is recomputed exactly on the plain path); the
    // queues get the storage their recorded peak needs (with a quarter of
    // headroom) when no source reached the reference capacity; the launch
    // shape (threads per block, source lanes per wave) is the fastest of a
    // few candidates, each tried once on successive builds.
    TuneRecord* tune = c.autotune ? c.tune : nullptr;
    std::vector<int32_t> order;
    // the predicted order is computed on a host thread while the device
    // memory is allocated; joined before the order is uploaded (or on return)
    struct Joiner {
        std::thread t;
        ~Joiner() { if (t.joinable()) t.join(); }
    } predictor;
    bool predicted = false;
    int blockSlot = -1;
    if (tune) {
        if (c.propBlock <= 0 && c.lanesPerWave <= 0) {
            int best = -1;
            for (int i = 0; i < kTuneShapes && blockSlot < 0; ++i) if (tune->blockMs[i] <= 0.0) blockSlot = i;
            if (blockSlot < 0) {
                for (int i = 0; i < kTuneShapes; ++i) if (best < 0 || tune->blockMs[i] < tune->blockMs[best]) best = i;
                blockSlot = best;
            }
            block = kTuneShape[blockSlot][0];
            lanesPerWave = kTuneShape[blockSlot][1];
        }
        if (c.queueCap <= 0 && (int)tune->cost.size() == V) {
            auto capOf = [](int peak) {
                int q = 1;
                while (q - 1 < peak + peak / 4 + 2) q <<= 1;
                return std::min(kHeapCap, q - 1);
            };
            if (tune->maxWinQ < kHeapCap) wcap = capOf(tune->maxWinQ);
            if (tune->maxPseudoQ < kHeapCap) pcap = capOf(tune->maxPseudoQ);
        }
    }
    // one block's LDS stays within 64 KiB
    while (block > 64 && lds_bytes(block, lanesPerWave) > (64u << 10)) block /= 2;
    const int lanes = c.lanes > 0 ? std::min(c.lanes, V) : std::min(V, resident_lanes(block, lanesPerWave));
    st.sourceOrder = 1;
    if (tune && (int)tune->cost.size() == V && lanes < V && c.sourceOrder == 0) {
        // Sources are taken from the costliest down (longest processing
        // time first): a lane runs about 1.25 sources on the torus, so a
        // costly source left to start late would set the end of the launch.
        costliest_first(tune->cost, order);
        st.sourceOrder = 3;
    } else if ((c.sourceOrder == 0 && lanes < V && V <= kPredictPerLane * lanes) || c.sourceOrder == 2) {
        try {
            predictor.t = std::thread([&h, K, &order, &predicted] {
                try {
                    predicted = predicted_order(h, K, order);
                } catch (...) {   // no memory for the prediction: id order
                    predicted = false;
                    order.clear();
                }
            });
        } catch (const std::system_error&) {
            // no thread to be had: the sources keep the id order
        }
    }
    if (tune && (int)tune->cost.size() == V) st.tuned = 1;
    if (vcap < 64) vcap = 64;
    if (scap < 64) scap = 64;
    // The extraction sorts a source's reached list in LDS (at most 8192
    // entries): larger tables are capped there, and a source that reaches
    // more vertices than the capped table holds is flagged and recomputed
    // on the plain path (large K).
    if (vcap > 8192) vcap = 8192;
    size_t freeB = 0, totB = 0;
    {
        const hipError_t me = hipMemGetInfo(&freeB, &totB);
        if (me != hipSuccess) return std::string("batched build: hipMemGetInfo: ") + hipGetErrorString(me);
    }
    // heap storage: the level 5-7 blocks (queues below 256 items) or all
    const int hslots = wcap < 256 ? kHeapSlotsMid : kHeapSlotsFull;
    const int phslots = pcap < 256 ? kHeapSlotsMid : kHeapSlotsFull;
    const int precUnits = (int)((sizeof(BPsRec) * (size_t)(pcap + 1) + sizeof(BWinRec) - 1) / sizeof(BWinRec));
    const size_t perLane = sizeof(BHeapEnt) * (size_t)(hslots + phslots) +
                           sizeof(BWinRec) * (size_t)(wcap + 1 + precUnits) + sizeof(uint16_t) * (size_t)(wcap + pcap + 2);
    size_t perSrc = sizeof(BSlotV) * vcap + sizeof(int32_t) * vcap + sizeof(BSlotS) * scap + 64;
    int chunk = c.chunk > 0 ? c.chunk : V;
    const size_t budget = freeB > ((size_t)lanes * perLane + (4ull << 30)) ? freeB - (size_t)lanes * perLane - (4ull << 30) : 0;
    if (st.tuned && (size_t)chunk * perSrc > budget && c.vertCap <= 0 && c.splitCap <= 0) {
        // the default tables would not all fit at once: size them to the
        // peak occupancy recorded (+3/4 load headroom) to keep one pass
        vcap = std::min(8192u, std::max(64u, next_pow2((unsigned)((tune->maxVert + 1) * 16 / 9 + 8))));
        scap = std::max(64u, next_pow2((unsigned)((tune->maxSplit + 1) * 16 / 9 + 8)));
        perSrc = sizeof(BSlotV) * vcap + sizeof(int32_t) * vcap + sizeof(BSlotS) * scap + 64;
    }
    {
        const size_t maxChunk = budget / perSrc;
        if ((size_t)chunk > maxChunk) chunk = (int)std::max<size_t>(maxChunk, 1);
    }
    if (chunk > V) chunk = V;
    st.lanes = lanes; st.propBlock = block; st.lanesPerWave = lanesPerWave; st.queueCap = wcap;
    st.pseudoCap = pcap; st.chunk = chunk; st.vertCap = (int)vcap; st.splitCap = (int)scap;

    auto t0 = std::chrono::steady_clock::now();
    BArgs a;
    std::memset(&a, 0, sizeof(a));
    a.m = d; a.g = g; a.K = K; a.lanes = lanes; a.vcap = vcap; a.scap = scap;
    a.vrefcap = vrefcap; a.srefcap = srefcap; a.hopCap = hop_cap(h);
    a.pseudoShare = kPseudoShare;
    a.wcap = wcap; a.pcap = pcap; a.hslots = hslots; a.phslots = phslots; a.precUnits = precUnits;
    a.lanesPerWave = lanesPerWave;
    std::string err;
    if (hipMalloc(&a.vtab, sizeof(BSlotV) * vcap * (size_t)chunk) != hipSuccess ||
        hipMalloc(&a.rlist, sizeof(int32_t) * vcap * (size_t)chunk) != hipSuccess ||
        hipMalloc(&a.stab, sizeof(BSlotS) * scap * (size_t)chunk) != hipSuccess ||
        hipMalloc(&a.wheap, sizeof(BHeapEnt) * (size_t)(hslots + phslots) * (size_t)lanes) != hipSuccess ||
        hipMalloc(&a.wrec, sizeof(BWinRec) * (size_t)(wcap + 1 + precUnits) * (size_t)lanes) != hipSuccess ||
        hipMalloc(&a.wfree, sizeof(uint16_t) * (size_t)(wcap + pcap + 2) * (size_t)lanes) != hipSuccess ||
        hipMalloc(&a.work, sizeof(int)) != hipSuccess ||
        hipMalloc(&a.status, sizeof(uint32_t) * (size_t)chunk) != hipSuccess ||
        hipMalloc(&a.vcount, sizeof(int32_t) * (size_t)chunk) != hipSuccess ||
        hipMalloc(&a.cnt, sizeof(int32_t) * kNumBatchedCounters * (size_t)chunk) != hipSuccess) {
        err = "batched build: out of device memory";
    }
    int32_t* dOrder = nullptr;
    if (predictor.t.joinable()) predictor.t.join();
    if (predicted) st.sourceOrder = 2;
    if (err.empty() && !order.empty()) {
        if (hipMalloc(&dOrder, sizeof(int32_t) * (size_t)V) != hipSuccess ||
            hipMemcpy(dOrder, order.data(), sizeof(int32_t) * (size_t)V, hipMemcpyHostToDevice) != hipSuccess)
            err = "batched build: out of device memory";
        a.order = dOrder;
    }
    std::vector<int32_t> newCost;
    int newMaxV = 0, newMaxS = 0, newMaxW = 0, newMaxP = 0;
    if (tune) newCost.assign((size_t)V, 0);
    st.allocSeconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::vector<int32_t> redo;
    st.status.assign((size_t)V, 0);
    st.counters.assign(CNT_NUM, 0);
    std::vector<uint32_t> hs;
    std::vector<int32_t> hc;
    hipEvent_t e0, e1, e2;
    hipEventCreate(&e0); hipEventCreate(&e1); hipEventCreate(&e2);
    for (int base = 0; err.empty() && base < V; base += chunk) {
        const int n = std::min(chunk, V - base);
        a.nsrc = n; a.srcBase = base;
        {
            hipError_t ce = hipMemsetAsync(a.vtab, 0xFF, sizeof(BSlotV) * vcap * (size_t)n, 0);
            if (ce == hipSuccess) ce = hipMemsetAsync(a.stab, 0xFF, sizeof(BSlotS) * scap * (size_t)n, 0);
            if (ce == hipSuccess) ce = hipMemsetAsync(a.work, 0, sizeof(int), 0);
            if (ce != hipSuccess) { err = std::string("batched build: table reset: ") + hipGetErrorString(ce); break; }
        }
        hipEventRecord(e0, 0);
        const int pl = std::min(lanes, n);
        a.lanes = pl;
        const size_t ldsBytes = lds_bytes(block, a.lanesPerWave);
        const size_t threads = (size_t)((pl + a.lanesPerWave - 1) / a.lanesPerWave) * 64;
        hipLaunchKernelGGL(batched_prop_kernel, dim3((unsigned)((threads + block - 1) / block)), dim3(block), ldsBytes,
                           0, a);
        hipEventRecord(e1, 0);
        const int eblocks = std::min(n, 16384);
        // an extracted source reached at most vrefcap vertices (beyond that
        // it is flagged and recomputed on the plain path)
        const unsigned sortCap = std::min(vcap, next_pow2((unsigned)std::max(vrefcap, 1)));
        if (sortCap <= 1024)
            hipLaunchKernelGGL(batched_extract_kernel<1024>, dim3(eblocks), dim3(64), 0, 0, a);
        else if (sortCap <= 2048)
            hipLaunchKernelGGL(batched_extract_kernel<2048>, dim3(eblocks), dim3(64), 0, 0, a);
        else if (sortCap <= 4096)
            hipLaunchKernelGGL(batched_extract_kernel<4096>, dim3(eblocks), dim3(64), 0, 0, a);
        else
            hipLaunchKernelGGL(batched_extract_kernel<8192>, dim3(eblocks), dim3(64), 0, 0, a);
        hipEventRecord(e2, 0);
        hipError_t ke = hipEventSynchronize(e2);
        if (ke == hipSuccess) ke = hipGetLastError();
        if (ke != hipSuccess) { err = std::string("batched kernels: ") + hipGetErrorString(ke); break; }
        const double pms = elapsed_ms(e0, e1), ems = elapsed_ms(e1, e2);
        st.propSeconds += pms * 1e-3; st.extractSeconds += ems * 1e-3; st.seconds += (pms + ems) * 1e-3;
        hs.resize((size_t)n); hc.resize((size_t)n * kNumBatchedCounters);
        {
            // the redo list is built from these flags: never from a copy that failed
            hipError_t ce = hipMemcpy(hs.data(), a.status, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToHost);
            if (ce == hipSuccess)
                ce = hipMemcpy(hc.data(), a.cnt, sizeof(int32_t) * kNumBatchedCounters * (size_t)n, hipMemcpyDeviceToHost);
            if (ce != hipSuccess) { err = std::string("batched build: status to host: ") + hipGetErrorString(ce); break; }
        }
        for (int i = 0; i < n; ++i) {
            const int src = order.empty() ? base + i : order[(size_t)base + i];
            st.status[src] = hs[i];
            if (hs[i] & (SF_CAPACITY | SF_REF_OVERFLOW)) redo.push_back(src);
            if (hs[i] & (SF_WIN_DROP | SF_PSEUDO_DROP)) st.flaggedDrop++;
            if (hs[i] & SF_REF_OVERFLOW) st.flaggedOverflow++;
            if (hs[i] & SF_ERROR) st.flaggedError++;
            const int32_t* cc = &hc[(size_t)i * kNumBatchedCounters];
            st.counters[CNT_WIN_PUSH] += cc[0]; st.counters[CNT_WIN_POP] += cc[1];
            st.counters[CNT_PSEUDO_PUSH] += cc[2]; st.counters[CNT_PSEUDO_POP] += cc[3];
            st.counters[CNT_PROPAGATE] += cc[4]; st.counters[CNT_CHILD_BUILT] += cc[5];
            st.counters[CNT_SPLIT_WRITE] += cc[6]; st.counters[CNT_VERT_WRITE] += cc[7];
            st.counters[CNT_FAN_HALFEDGE] += cc[8]; st.counters[CNT_PSEUDO_NEIGH] += cc[9];
            st.counters[CNT_BT_HOPS] += cc[10]; st.counters[CNT_REACHED] += cc[11];
            if (tune) {
                newCost[src] = cc[1] + kPseudoCost * cc[3];
                // a source that hit the capacity stopped early: its peak is unknown
                const int grow = (hs[i] & SF_CAPACITY) ? 2 : 1;
                newMaxV = std::max(newMaxV, cc[11] * grow);
                newMaxS = std::max(newMaxS, cc[12] * grow);
                // a source stopped by its queue storage has an unknown peak: use the reference's
                newMaxW = std::max(newMaxW, (hs[i] & SF_CAPACITY) ? kHeapCap : cc[13]);
                newMaxP = std::max(newMaxP, (hs[i] & SF_CAPACITY) ? kHeapCap : cc[14]);
            }
        }
    }
    if (tune && err.empty()) {
        tune->cost.swap(newCost);
        tune->maxVert = newMaxV;
        tune->maxSplit = newMaxS;
        tune->maxWinQ = newMaxW;
        tune->maxPseudoQ = newMaxP;
        if (blockSlot >= 0) tune->blockMs[blockSlot] = st.propSeconds * 1e3;
    }
    hipFree(dOrder);
    hipEventDestroy(e0); hipEventDestroy(e1); hipEventDestroy(e2);
    hipFree(a.vtab); hipFree(a.rlist); hipFree(a.stab); hipFree(a.wheap); hipFree(a.wrec);
    hipFree(a.wfree); hipFree(a.work); hipFree(a.status); hipFree(a.vcount); hipFree(a.cnt);
    if (!err.empty()) return err;
    if (!redo.empty()) {
        // Exact recomputation on the reference-layout path (its own tables
        // replicate the reference, including overflow aliasing).
        BuildConfig pc = c;
        pc.plainBlocks = std::max(1, std::min(512, (int)((redo.size() + 127) / 128)));
        pc.plainThreads = 128;
        BuildStats ps;
        if (c.ensurePlainMesh) {
            const std::string ue = c.ensurePlainMesh();
            if (!ue.empty()) return "redo of flagged sources: " + ue;
        }
        err = build_plain(h, d, pc, g, ps, redo.data(), (int)redo.size());
        if (!err.empty()) return "redo of flagged sources: " + err;
        for (int v : redo) st.status[v] = ps.status[v] | SF_REDONE | (st.status[v] & (SF_CAPACITY | SF_REF_OVERFLOW));
        st.redoSeconds += ps.redoSeconds;
        st.seconds += ps.redoSeconds;
        st.redone = (int)redo.size();
    }
    return std::string();
}

}  // namespace svg
