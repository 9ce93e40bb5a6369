// Reference-layout ICH storage for one source per thread: the reference's
// records (80-B window heap items, 40-B pseudo-source items), MyHashTable
// with its hash, linear probing and overflow slot, and Clear() of every slot.
// Used by the plain build kernel and by the surface-point queries.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "ich_engine.h"
#include "svg_graph.h"
#include "svg_mesh.h"

namespace svg {

struct PWin {                   // ICHDevice::Window (ICHDevice.cuh:16-24), 72 B
    unsigned edgeID;
    double b0, b1, d0, d1, ps, minDist;
    unsigned srcID, pseudoSrcId;
    int birth;
    int level;
};
struct PWinItem { double key; PWin item; };          // PQItem<Window>, 80 B
struct PPs {                    // ICHDevice::PseudoWindow, 32 B
    unsigned vertID;
    double dist;
    unsigned srcId, pseudoSrcId;
    unsigned birth;
    unsigned level;
};
struct PPsItem { double key; PPs item; };            // PQItem<PseudoWindow>, 40 B
struct PSplitItem { double dist, x; unsigned ps; int index; };
struct PVertItem {
    int8_t birth;
    double dist;
    uint8_t isSource;
    int enter;
    unsigned ps, srcId;
    int index;
};

struct PlainPolicy {
    static constexpr bool kFanRecords = false;
    const EdgeAoS* __restrict__ E;
    const VertAoS* __restrict__ Vt;
    const double* __restrict__ cA;
    const double* __restrict__ cB;
    int nv_, nf_;
    PSplitItem* sp; unsigned ssz;
    PVertItem* vt; unsigned vsz;
    PWinItem* wq; int wtail;
    PPsItem* pq; int ptail;
    uint32_t flags;
    int srcVert;

    __device__ int twin(int e) const { return E[e].twin; }
    __device__ int next(int e) const { return E[e].next; }
    __device__ int prev(int e) const { return E[e].prev; }
    __device__ int v0(int e) const { return E[e].v0; }
    __device__ int v1(int e) const { return E[e].v1; }
    __device__ double len(int e) const { return E[e].len; }
    __device__ int first_edge(int v) const { return Vt[v].firstEdge; }
    __device__ uint8_t vflag(int v) const { return (uint8_t)Vt[v].flags; }
    __device__ double cornerA(int e) const { return cA[e]; }
    __device__ double cornerB(int e) const { return cB[e]; }
    __device__ int nv() const { return nv_; }
    __device__ int nf() const { return nf_; }
    __device__ bool stop() const { return false; }
    __device__ void face3(int e0, Tri& t) const
    {
        t.e0 = e0; t.e1 = E[e0].next; t.e2 = E[t.e1].next;
        t.a = E[e0].v0; t.b = E[e0].v1; t.op = E[t.e1].v1;
        t.l0 = E[e0].len; t.l1 = E[t.e1].len; t.l2 = E[t.e2].len;
        t.tw1 = E[t.e1].twin; t.tw2 = E[t.e2].twin;
    }
    // event counts are not kept on the reference-layout path; an error
    // (a source ended on unordered queue keys) still sets its status bit
    __device__ void count(int c, long long) { if (c == CNT_ERRORS) flags |= SF_ERROR; }

    // MyHashTable (MyHashTable.cuh:98-117)
    __device__ static unsigned hfun(unsigned index, unsigned size)
    {
        double p = index * 0.618033989;
        return (unsigned)(int)((p - (int)p) * size);
    }
    __device__ unsigned vkey(unsigned idx)
    {
        unsigned k = hfun(idx, vsz), i = 0;
        for (; i < vsz; ++i) {
            int ix = vt[k].index;
            if ((unsigned)ix == idx || ix == -1) break;
            k = (k + 1) % vsz;
        }
        if (i == vsz) { k = vsz; flags |= SF_REF_OVERFLOW; }
        return k;
    }
    __device__ unsigned skey(unsigned idx)
    {
        unsigned k = hfun(idx, ssz), i = 0;
        for (; i < ssz; ++i) {
            int ix = sp[k].index;
            if ((unsigned)ix == idx || ix == -1) break;
            k = (k + 1) % ssz;
        }
        if (i == ssz) { k = ssz; flags |= SF_REF_OVERFLOW; }
        return k;
    }
    __device__ VInfo vget(int v)
    {
        const PVertItem& it = vt[vkey((unsigned)v)];
        VInfo r; r.dist = it.dist; r.enter = it.enter; r.birth = it.birth; r.isSource = it.isSource;
        return r;
    }
    __device__ void vput(int v, const VInfo& x)
    {
        unsigned k = vkey((unsigned)v);
        vt[k].index = v;
        vt[k].birth = x.birth; vt[k].dist = x.dist; vt[k].enter = x.enter; vt[k].isSource = x.isSource;
        vt[k].srcId = (unsigned)srcVert;
    }
    __device__ void vtouch(int v) { unsigned k = vkey((unsigned)v); vt[k].index = v; }
    __device__ void push_hint() {}
    __device__ int birth_of(int v) { return (int)vget(v).birth; }
    __device__ void lookup_face(const Tri& t, VInfo& va, VInfo& vb, VInfo& vo, SInfo& s)
    {
        va = vget(t.a); vb = vget(t.b); vo = vget(t.op); s = sget(t.e0);
    }
    __device__ SInfo sget(int e)
    {
        const PSplitItem& it = sp[skey((unsigned)e)];
        SInfo r; r.dist = it.dist; r.x = it.x; r.ps = (int)it.ps;
        return r;
    }
    __device__ void sput(int e, const SInfo& x)
    {
        unsigned k = skey((unsigned)e);
        sp[k].index = e;
        sp[k].dist = x.dist; sp[k].x = x.x; sp[k].ps = (unsigned)x.ps;
    }

    // PriorityQueues (PriorityQueue.cuh:51-105)
    __device__ bool wempty() const { return wtail == 1; }
    __device__ void wtop(Win& w) const
    {
        const PWin& it = wq[1].item;
        w.edge = (int)it.edgeID; w.b0 = it.b0; w.b1 = it.b1; w.d0 = it.d0; w.d1 = it.d1; w.ps = it.ps;
        win_finish(w);
        w.twin = E[w.edge].twin;
        w.pseudoSrc = (int)it.pseudoSrcId; w.birth = it.birth; w.level = it.level;
    }
    __device__ void wpop()
    {
        --wtail;
        const PWinItem top = wq[wtail];
        int cur = 1, l = 2, r = 3, mc = -1;
        if (r < wtail) mc = wq[l].key < wq[r].key ? l : r;
        else if (l < wtail) mc = l;
        while (mc != -1 && wq[mc].key < top.key) {
            wq[cur] = wq[mc];
            cur = mc; l = cur * 2; r = cur * 2 + 1;
            if (r < wtail) mc = wq[l].key < wq[r].key ? l : r;
            else if (l < wtail) mc = l;
            else mc = -1;
        }
        wq[cur] = top;
    }
    __device__ void wpush(const Win& w)
    {
        if (wtail - 1 >= kHeapCap) { flags |= SF_WIN_DROP; return; }
        ++wtail;
        int cur = wtail - 1, p = cur / 2;
        const double key = w.minDist;
        while (p > 0 && wq[p].key > key) { wq[cur] = wq[p]; cur = p; p = cur / 2; }
        PWin it;
        it.edgeID = (unsigned)w.edge; it.b0 = w.b0; it.b1 = w.b1; it.d0 = w.d0; it.d1 = w.d1;
        it.ps = w.ps; it.minDist = w.minDist; it.srcID = (unsigned)srcVert;
        it.pseudoSrcId = (unsigned)w.pseudoSrc; it.birth = w.birth; it.level = w.level;
        wq[cur].item = it;
        wq[cur].key = key;
    }
    __device__ bool pempty() const { return ptail == 1; }
    __device__ void ptop(Pseudo& q) const
    {
        const PPs& it = pq[1].item;
        q.vert = (int)it.vertID; q.dist = it.dist; q.pseudoSrc = (int)it.pseudoSrcId;
        q.birth = (int)it.birth; q.level = (int)it.level;
    }
    __device__ void ppop()
    {
        --ptail;
        const PPsItem top = pq[ptail];
        int cur = 1, l = 2, r = 3, mc = -1;
        if (r < ptail) mc = pq[l].key < pq[r].key ? l : r;
        else if (l < ptail) mc = l;
        while (mc != -1 && pq[mc].key < top.key) {
            pq[cur] = pq[mc];
            cur = mc; l = cur * 2; r = cur * 2 + 1;
            if (r < ptail) mc = pq[l].key < pq[r].key ? l : r;
            else if (l < ptail) mc = l;
            else mc = -1;
        }
        pq[cur] = top;
    }
    __device__ void ppush(const Pseudo& q)
    {
        if (ptail - 1 >= kHeapCap) { flags |= SF_PSEUDO_DROP; return; }
        ++ptail;
        int cur = ptail - 1, p = cur / 2;
        const double key = q.dist;
        while (p > 0 && pq[p].key > key) { pq[cur] = pq[p]; cur = p; p = cur / 2; }
        PPs it;
        it.vertID = (unsigned)q.vert; it.dist = q.dist; it.srcId = (unsigned)srcVert;
        it.pseudoSrcId = (unsigned)q.pseudoSrc; it.birth = (unsigned)q.birth; it.level = (unsigned)q.level;
        pq[cur].item = it;
        pq[cur].key = key;
    }

    // ICHDevice::Clear (ICHDevice.cu:487-523): every slot back to defaults.
    __device__ void clear()
    {
        wtail = 1; ptail = 1; flags = 0;
        for (unsigned i = 0; i < ssz + 1; ++i) {
            sp[i].dist = DBL_MAX; sp[i].x = DBL_MAX; sp[i].ps = 0xFFFFFFFFu; sp[i].index = -1;
        }
        for (unsigned i = 0; i < vsz + 1; ++i) {
            vt[i].birth = -1; vt[i].dist = DBL_MAX; vt[i].enter = -1; vt[i].isSource = 0;
            vt[i].ps = 0xFFFFFFFFu; vt[i].srcId = 0xFFFFFFFFu; vt[i].index = -1;
        }
    }
};

}  // namespace svg
