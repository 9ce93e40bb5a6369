// Multi-query bucketed frontier engine (see solve.h) and the query entry
// points of the C ABI (svg_sssd, svg_msad, svg_sssd_batch).
//
// A round is three passes: classify (frontier entries -> deferred list and
// rows to relax), expand (relax the rows, push improved targets) and
// bookkeep (per-slot bucket bounds).  Device-scope atomics on MI355X are
// performed at the memory side, where atomics to one address serialise
// (about 11 ns each), so no pass funnels its output through a single
// counter: classify writes into static per-block regions sized by its input
// range, expand buffers its pushes per wave in LDS and appends them to one of
// kPLists lists, and per-slot statistics are spread over kAccCopies copies.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "point.h"
#include "solve.h"
#include "svg_capi_internal.h"

namespace svg {

constexpr unsigned long long kInfBits = 0x7FEFFFFFFFFFFFFFull;   // DBL_MAX (GraphDistInfo default)
constexpr int kSolveBlock = 256;
constexpr int kSolveWaves = kSolveBlock / 64;
constexpr int kPBuf = 512;              // per-wave push buffer (entries)
constexpr int kChunk = 64;              // won-list chunk
constexpr uint64_t kSentinel = ~0ull;
constexpr int kSegs = kCBlocks + kPLists;   // classify input segments

struct SolveArgs {
    int V, K;
    const int32_t* __restrict__ tail;
    const int32_t* __restrict__ adj;
    const double* __restrict__ w;
    const double* __restrict__ pos;
    unsigned long long* dist;
    int32_t* parent;
    uint32_t* stamp;
    uint32_t* inq;
    int32_t* srcid;
    const float* __restrict__ hdist;    // S*V straight-line distance to the SSSD target, rounded down
    // lists of this round (c = input parity)
    const uint64_t* dIn;                // deferred regions read by classify
    uint64_t* dOut;                     // deferred regions written by classify
    const uint64_t* pIn;                // kPLists push lists read by classify
    uint64_t* pOut;                     // kPLists push lists written by expand
    ExpEnt* exp;                        // rows to relax this round (classify regions)
    const ExpEnt* expPrev;              // rows of the previous round
    SolveCounts* cnt;
    int c;                              // parity of this round
    int doBook;                         // classify bookkeeps the previous round (S <= kLdsSlots)
    int doParent;                       // classify resolves the previous round's parents
    size_t frontCap;                    // capacity of a deferred / row list
    size_t pCap;                        // capacity of one push list
    SlotParam* sp;
    SlotAcc* acc;                       // kAccCopies * S, this round's statistics
    SlotAcc* accPrev;                   // the previous round's
    double* thr;                        // 2 * S bucket bounds by parity
    int* doneP;                         // 2 * S finished flags by parity
    int S;
    uint32_t round;
    WonEnt* won;                        // this round's won lists
    const WonEnt* wonPrev;              // the previous round's
    size_t wonCap;                      // per won list
    int* ovFlag;                        // set when a list would overflow
    int chainBudget;                    // rows a wave may chain per expand pass
};

// A join slot's bound (SlotParam::result), lowered by atomicMin on its bits
// while expand passes run.
__device__ __forceinline__ double join_bound(const SlotParam& sp)
{
    const unsigned long long b =
        __hip_atomic_load((const unsigned long long*)&sp.result, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __longlong_as_double((long long)b);
}

__device__ __forceinline__ double heur(const double* __restrict__ pos, int v, int t)
{
    // (verts[v].pos - verts[t].pos).length()  (SVG.cu:402)
    const double x = pos[4 * v] - pos[4 * t];
    const double y = pos[4 * v + 1] - pos[4 * t + 1];
    const double z = pos[4 * v + 2] - pos[4 * t + 2];
    return sqrt(x * x + y * y + z * z);
}

__device__ __forceinline__ uint64_t pack_entry(int gen, int slot, int v)
{
    return ((uint64_t)(uint16_t)gen << 48) | ((uint64_t)(uint16_t)slot << 32) | (uint64_t)(uint32_t)v;
}

__device__ __forceinline__ unsigned long long lanes_below(int lane)
{
    return (1ull << lane) - 1ull;
}

// Exclusive prefix over n counts into pre[0..n] (pre[n] = total), by one
// wave; the caller synchronises the block afterwards.
__device__ void wave_prefix(const int* cnt, int n, int* pre, int lane)
{
    const int per = (n + 63) / 64;
    const int lo = lane * per, hi = min(n, lo + per);
    int s = 0;
    for (int i = lo; i < hi; ++i) s += cnt[i];
    int incl = s;
    for (int off = 1; off < 64; off <<= 1) {
        const int o = __shfl_up(incl, off);
        if (lane >= off) incl += o;
    }
    int run = incl - s;
    for (int i = lo; i < hi; ++i) { pre[i] = run; run += cnt[i]; }
    if (lane == 63) pre[n] = incl;
}

// Index of the segment holding element g: the last s with pre[s] <= g.
__device__ __forceinline__ int seg_of(const int* pre, int n, int g)
{
    int lo = 0, hi = n;                 // pre[lo] <= g < pre[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pre[mid] <= g) lo = mid; else hi = mid;
    }
    return lo;
}

// Resets the admitted slots and, for an A* slot, tabulates the straight-line
// distance of every vertex to its target (the heuristic of SVG.cu:402).
__global__ void k_slot_init(unsigned long long* dist, int32_t* parent, uint32_t* stamp, uint32_t* inq,
                            int32_t* srcid, float* hdist, const double* __restrict__ pos,
                            const int* slots, const int* targets, int nslots, int V)
{
    const size_t total = (size_t)nslots * (size_t)V;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(i / V);
        const int v = (int)(i - (size_t)k * V);
        const size_t o = (size_t)slots[k] * V + v;
        dist[o] = kInfBits;
        parent[o] = -1;
        stamp[o] = 0;
        inq[o] = 0;
        srcid[o] = -1;
        const int t = targets[k];
        // rounded toward zero, it stays a lower bound (admissible) in 4 bytes;
        // a join slot (t = -2) starts with no exits
        hdist[o] = t >= 0 ? __double2float_rz(heur(pos, v, t)) : (t == -2 ? __int_as_float(0x7f800000) : 0.0f);
    }
}

// Bookkeeping of one slot after a round with statistics `acc` (kAccCopies
// copies): with nothing left below the bound the bound moves to the smallest
// key + delta, or the slot finishes.  Keys of the next round's entries are
// the deferred ones (at least minDeferred, or lower if relaxed since, which
// minImp covers) and the relaxed targets (at least minImp), so when none is
// below the bound the next round would expand nothing and the bound moves
// now instead of a round later.  With `write` the slot's record is updated.
__device__ void book_slot(const SolveArgs& a, int q, const SlotAcc* acc, double thrOld, int doneOld,
                          double& thrNew, int& doneNew, bool write)
{
    unsigned long long md = ~0ull, mi = ~0ull, settled = 0, relaxed = 0;
    int expanded = 0;
    for (int k = 0; k < kAccCopies; ++k) {
        const SlotAcc& x = acc[(size_t)k * a.S + q];
        md = x.minDeferred < md ? x.minDeferred : md;
        mi = x.minImp < mi ? x.minImp : mi;
        settled += x.settled;
        relaxed += x.relaxed;
        expanded |= x.expanded;
    }
    thrNew = thrOld;
    doneNew = doneOld;
    SlotParam& sp = a.sp[q];
    if (sp.kind < 0 || doneOld) return;
    const unsigned long long nextMin = md < mi ? md : mi;
    const double nm = __longlong_as_double((long long)nextMin);
    if (nextMin == ~0ull) {
        doneNew = 1;
    } else if (!(nm < thrOld)) {
        bool finished = false;
        if (sp.kind == Q_SSSD) {
            const double ub = __longlong_as_double((long long)a.dist[(size_t)q * a.V + sp.target]);
            finished = nm * (1.0 - 1e-12) > ub;
        } else if (sp.kind == Q_JOIN) {
            // every vertex still to settle is at least nm away: its exit sum
            // exceeds the bound (which is at least the exact best sum)
            finished = nm * (1.0 - 1e-12) > join_bound(sp);
        }
        if (finished) doneNew = 1;
        else thrNew = nm + sp.delta;
    }
    if (write) {
        sp.rounds++;
        sp.settled += (long long)settled;
        sp.relaxed += (long long)relaxed;
        sp.expanded = expanded;
        sp.minDeferred = md;
        sp.thr = thrNew;
        sp.done = doneNew;
        if (sp.kind == Q_SSSD) sp.result = __longlong_as_double((long long)a.dist[(size_t)q * a.V + sp.target]);
    }
}

// Parents of the vertices improved in round `rnd` (parity p): the smallest u
// whose expansion value reproduces the new distance exactly.  The round's
// improvements and exact ties are the only candidates; if a won list
// overflowed, the round's rows are scanned again instead.  `pre` is LDS of
// kCBlocks + 1 ints; blocks vb = 0 .. nvb-1 of the grid call this.
__device__ void parent_part(const SolveArgs& a, int p, uint32_t rnd, const WonEnt* won, const ExpEnt* rows,
                            int* pre, int vb, int nvb)
{
    SolveCounts* C = a.cnt;
    int cnt[kPLists];
    bool over = false;
    int total = 0;
#pragma unroll
    for (int l = 0; l < kPLists; ++l) {
        cnt[l] = C->wcnt[p][l * kCntPad];
        over |= (size_t)cnt[l] > a.wonCap;
        total += cnt[l];
    }
    if (!over) {
        // one pass over the concatenated lists
        for (int i = vb * blockDim.x + threadIdx.x; i < total; i += nvb * blockDim.x) {
            int l = 0, off = i;
            while (off >= cnt[l]) { off -= cnt[l]; ++l; }
            const WonEnt we = won[(size_t)l * a.wonCap + off];
            if (we.slot < 0) continue;
            const size_t o = (size_t)we.slot * a.V + we.x;
            if (a.stamp[o] == rnd && a.dist[o] == we.nb) atomicMin(&a.parent[o], we.u);
        }
        return;
    }
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    if (wid == 0) wave_prefix(C->ecnt[p], kCBlocks, pre, lane);
    __syncthreads();
    const int n = pre[kCBlocks];
    const int stride = C->estride[p];
    const int nw = nvb * kSolveWaves;
    for (int g = vb * kSolveWaves + wid; g < n; g += nw) {
        const int sgi = seg_of(pre, kCBlocks, g);
        const ExpEnt ex = rows[(size_t)sgi * stride + (g - pre[sgi])];
        const int qq = (int)((ex.entry >> 32) & 0xFFFFu);
        if (!a.sp[qq].wantParent) continue;
        const int vv = (int)(ex.entry & 0xFFFFFFFFu);
        const double dvv = __longlong_as_double((long long)ex.dist);
        const size_t sb = (size_t)qq * a.V;
        const int tl = a.tail[vv];
        const size_t row = (size_t)vv * a.K;
        for (int i = lane; i < tl; i += 64) {
            const int x = a.adj[row + i];
            if (a.stamp[sb + x] != rnd) continue;
            const double nd = dvv + a.w[row + i];
            if ((unsigned long long)__double_as_longlong(nd) == a.dist[sb + x])
                atomicMin(&a.parent[sb + x], vv);
        }
    }
    __syncthreads();
}

// Pass 1: the input (deferred regions of the previous classify, then the
// push lists) is cut into kCBlocks equal ranges, one per block.  A block's
// outputs fit in the same range of the output lists, so they are placed with
// LDS counters only.  No other pass runs concurrently, so frontier flags and
// distances are read and written plainly: launch boundaries order them
// against the relaxations of pass 2.
__global__ void __launch_bounds__(kSolveBlock) k_classify(SolveArgs a)
{
    __shared__ int cnts[kSegs];
    __shared__ int pre[kSegs + 1];
    __shared__ int sD, sE;
    // per-slot table (up to kLdsSlots slots): bound, state, generation and,
    // for A*, the current answer (no distance changes during this pass)
    __shared__ double lThr[kLdsSlots], lUb[kLdsSlots];
    __shared__ int lDone[kLdsSlots], lKind[kLdsSlots], lGen[kLdsSlots];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    SolveCounts* C = a.cnt;
    const bool table = a.S <= kLdsSlots;
    // With doBook, the previous round's bookkeeping, computed identically by
    // every block (block 0 records it); then the previous round's parents.
    if (table) {
        for (int q = threadIdx.x; q < a.S; q += blockDim.x) {
            const SlotParam& sp = a.sp[q];
            double t; int d;
            if (a.doBook) {
                book_slot(a, q, a.accPrev, a.thr[(a.c ^ 1) * a.S + q], a.doneP[(a.c ^ 1) * a.S + q], t, d,
                          blockIdx.x == 0);
                if (blockIdx.x == 0) { a.thr[a.c * a.S + q] = t; a.doneP[a.c * a.S + q] = d; }
            } else {
                t = a.thr[a.c * a.S + q];
                d = a.doneP[a.c * a.S + q];
            }
            const int kind = sp.kind;
            lThr[q] = t; lDone[q] = d; lKind[q] = kind; lGen[q] = sp.gen;
            lUb[q] = kind == Q_SSSD ? __longlong_as_double((long long)a.dist[(size_t)q * a.V + sp.target])
                                    : (kind == Q_JOIN ? join_bound(sp) : 0.0);
        }
    }
    if (a.doParent) parent_part(a, a.c ^ 1, a.round - 1, a.wonPrev, a.expPrev, pre, blockIdx.x, gridDim.x);
    // a push list that overflowed (flagged for the host) is read up to its capacity
    for (int i = threadIdx.x; i < kSegs; i += blockDim.x)
        cnts[i] = i < kCBlocks ? C->dcnt[a.c][i]
                               : (int)min((size_t)C->pcnt[a.c][(i - kCBlocks) * kCntPad], a.pCap);
    if (threadIdx.x == 0) { sD = 0; sE = 0; }
    __syncthreads();
    if (wid == 0) wave_prefix(cnts, kSegs, pre, lane);
    __syncthreads();
    int N = pre[kSegs];
    if ((size_t)N + kCBlocks > a.frontCap) {       // cannot happen while the lists hold one
        N = (int)(a.frontCap - kCBlocks);          // entry per (slot, vertex) and query
        if (threadIdx.x == 0) atomicOr(a.ovFlag, 1);
    }
    const int R = (N + kCBlocks - 1) / kCBlocks;
    const int dStrideIn = C->dstride[a.c];
    const int b = blockIdx.x;
    if (b == 0 && threadIdx.x == 0) { C->dstride[a.c ^ 1] = R; C->estride[a.c] = R; }
    const int g1 = min(N, (b + 1) * R);
    const size_t outBase = (size_t)b * R;
    SlotAcc* acc = a.acc + (size_t)(b % kAccCopies) * a.S;
    for (int g0 = b * R + wid * 64; g0 < g1; g0 += kSolveBlock) {
        const int g = g0 + lane;
        uint64_t e = kSentinel;
        if (g < g1) {
            const int s = seg_of(pre, kSegs, g);
            const int off = g - pre[s];
            e = s < kCBlocks ? a.dIn[(size_t)s * dStrideIn + off] : a.pIn[(size_t)(s - kCBlocks) * a.pCap + off];
        }
        int q = 0, act = 0;           // act: 0 none, 1 defer, 2 expand
        unsigned long long defKey = ~0ull, db = 0;
        int rowLen = 0;
        if (e != kSentinel) {
            q = (int)((e >> 32) & 0xFFFFu);
            const int v = (int)(e & 0xFFFFFFFFu);
            const size_t o = (size_t)q * a.V + v;
            const int tlv = a.tail[v];       // loaded with the entry's other reads
            double thr, ub = 0.0;
            int done, kind, gen;
            if (table) {
                thr = lThr[q]; done = lDone[q]; kind = lKind[q]; gen = lGen[q]; ub = lUb[q];
            } else {
                const SlotParam& sp = a.sp[q];
                thr = a.thr[a.c * a.S + q]; done = a.doneP[a.c * a.S + q]; kind = sp.kind; gen = sp.gen;
                if (kind == Q_SSSD) ub = __longlong_as_double((long long)a.dist[(size_t)q * a.V + sp.target]);
                else if (kind == Q_JOIN) ub = join_bound(sp);
            }
            if (kind < 0 || (int)(e >> 48) != (gen & 0xFFFF)) {
                // entry of an earlier query in this slot: dropped
            } else if (done) {
                a.inq[o] = 0;
            } else {
                db = a.dist[o];
                const double dv = __longlong_as_double((long long)db);
                double key = dv;
                bool prune = false;
                if (kind == Q_SSSD) {
                    key = dv + (double)a.hdist[o];
                    prune = key * (1.0 - 1e-12) > ub;
                } else if (kind == Q_JOIN) {
                    prune = dv * (1.0 - 1e-12) > ub;
                }
                if (prune) {
                    // no path through v can beat the answer any more
                    a.inq[o] = 0;
                } else if (key >= thr) {
                    // deferred: v stays flagged and moves to the next frontier
                    act = 1;
                    defKey = (unsigned long long)__double_as_longlong(key);
                } else {
                    act = 2;
                    a.inq[o] = 0;
                    rowLen = tlv;
                }
            }
        }
        const unsigned long long dm = __ballot(act == 1);
        if (dm) {
            int base = 0;
            if (lane == 0) base = atomicAdd(&sD, __popcll(dm));
            base = __shfl(base, 0);
            if (act == 1) a.dOut[outBase + base + __popcll(dm & lanes_below(lane))] = e;
        }
        const unsigned long long em = __ballot(act == 2);
        if (em) {
            int base = 0;
            if (lane == 0) base = atomicAdd(&sE, __popcll(em));
            base = __shfl(base, 0);
            if (act == 2) {
                // the row length replaces the generation tag (rows are checked)
                ExpEnt x; x.entry = (e & 0x0000FFFFFFFFFFFFull) | ((uint64_t)(uint32_t)rowLen << 48); x.dist = db;
                a.exp[outBase + base + __popcll(em & lanes_below(lane))] = x;
            }
        }
        // one update per slot present in this wave, into this block's copy
        unsigned long long todo = __ballot(act != 0);
        while (todo) {
            const int leader = __ffsll((long long)todo) - 1;
            const int lq = __shfl(q, leader);
            const bool mine = act != 0 && q == lq;
            todo &= ~__ballot(mine);
            unsigned long long k = mine ? defKey : ~0ull;
            long long arcs = mine ? rowLen : 0;
            for (int off = 32; off > 0; off >>= 1) {
                const unsigned long long o = __shfl_xor(k, off);
                k = o < k ? o : k;
                arcs += __shfl_xor(arcs, off);
            }
            const int nExp = __popcll(__ballot(mine && act == 2));
            if (lane == leader) {
                if (k != ~0ull) atomicMin(&acc[lq].minDeferred, k);
                if (nExp) {
                    atomicOr(&acc[lq].expanded, 1);
                    atomicAdd(&acc[lq].settled, (unsigned long long)nExp);
                    atomicAdd(&acc[lq].relaxed, (unsigned long long)arcs);
                }
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) { C->dcnt[a.c ^ 1][b] = sD; C->ecnt[a.c][b] = sE; }
}

// Appends the wave's buffered pushes to push list `list` with one atomic.
__device__ __forceinline__ void flush_pushes(const SolveArgs& a, uint64_t* buf, int& n, int list, int lane)
{
    if (n == 0) return;
    int base = 0;
    if (lane == 0) base = atomicAdd(&a.cnt->pcnt[a.c ^ 1][list * kCntPad], n);
    base = __shfl(base, 0);
    uint64_t* dst = a.pOut + (size_t)list * a.pCap;
    for (int j = lane; j < n; j += 64) {
        if ((size_t)(base + j) < a.pCap) dst[base + j] = buf[j];
        else atomicOr(a.ovFlag, 1);
    }
    __builtin_amdgcn_wave_barrier();
    n = 0;
}

// Reserves room for the lanes of `mask` in the wave's won-list chunk.
__device__ __forceinline__ int won_pos(const SolveArgs& a, WonEnt* list, int* counter, int& base, int& used,
                                       unsigned long long mask, int lane)
{
    const int cnt = __popcll(mask);
    if (used + cnt > kChunk) {
        for (int j = used + lane; j < kChunk; j += 64)
            if ((size_t)(base + j) < a.wonCap) list[base + j].slot = -1;
        int nb = 0;
        if (lane == 0) nb = atomicAdd(counter, kChunk);
        base = __shfl(nb, 0);
        used = 0;
    }
    const int pos = base + used + __popcll(mask & lanes_below(lane));
    used += cnt;
    return pos;
}

// Pass 2: one wave per row, rows dealt round-robin over all waves.  Up to
// kSteps 64-arc steps are processed in phases so each phase's memory
// operations are in flight together: arc loads, target-distance loads, the
// A* bound's loads, the atomicMins, the frontier-flag exchanges.  The plain
// read of a target's distance filters the relaxations first; it can only be
// stale-high (distances only fall), so no improvement is missed.
//
// Chaining: a target this wave improves, whose frontier flag it takes and
// whose key is still inside the slot's bucket, would be expanded by the next
// round anyway; up to a budget of such rows per wave are expanded by the
// wave itself before the pass ends, so one round advances the frontier by
// several hops.  Any relaxation order reaches the same least fixed point,
// and a chained row reads its distance after clearing its flag (both where
// the atomics are performed), so an improvement that lands later pushes the
// vertex again.
constexpr int kSteps = 8;
constexpr int kChainCap = 64;           // per-wave queue of chained rows (LDS)

__global__ void __launch_bounds__(kSolveBlock) k_expand(SolveArgs a)
{
    __shared__ int pre[kCBlocks + 1];
    __shared__ uint64_t pbuf[kSolveWaves][kPBuf];
    __shared__ int4 lSlot[kLdsSlots];           // kind, target, wantParent, gen
    __shared__ double lThr[kLdsSlots];          // bucket bound of this round
    __shared__ int lDone[kLdsSlots];
    __shared__ int2 cq[kSolveWaves][kChainCap]; // chained rows: slot, vertex
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    SolveCounts* C = a.cnt;
    const bool table = a.S <= kLdsSlots;
    if (table)
        for (int q = threadIdx.x; q < a.S; q += blockDim.x) {
            const SlotParam& sp = a.sp[q];
            lSlot[q] = make_int4(sp.kind, sp.target, sp.wantParent, sp.gen);
            lThr[q] = a.thr[a.c * a.S + q];
            lDone[q] = a.doneP[a.c * a.S + q];
        }
    // state the classify pass has consumed is cleared for the next round
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < kAccCopies * a.S; i += gridDim.x * blockDim.x) {
        SlotAcc& x = a.accPrev[i];
        x.minDeferred = ~0ull; x.minImp = ~0ull; x.settled = 0; x.relaxed = 0; x.expanded = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x < kPLists) {
        C->pcnt[a.c][threadIdx.x * kCntPad] = 0;
        C->wcnt[a.c ^ 1][threadIdx.x * kCntPad] = 0;
    }
    if (wid == 0) wave_prefix(C->ecnt[a.c], kCBlocks, pre, lane);
    __syncthreads();
    const int n = pre[kCBlocks];
    const int stride = C->estride[a.c];
    const int wave = blockIdx.x * kSolveWaves + wid;
    const int nw = gridDim.x * kSolveWaves;
    const int list = blockIdx.x % kPLists;
    uint64_t* buf = pbuf[wid];
    int nBuf = 0;
    WonEnt* wonL = a.won + (size_t)list * a.wonCap;
    int* wonC = &C->wcnt[a.c][list * kCntPad];
    int wBase = 0, wUsed = kChunk;
    SlotAcc* acc = a.acc + (size_t)(blockIdx.x % kAccCopies) * a.S;
    const int chainMax = table ? min(a.chainBudget, kChainCap) : 0;
    int nChain = 0;

    // relaxes the arcs of row vv of slot qq from distance dvv
    auto relax_row = [&](int qq, int vv, double dvv, int tl) {
        int4 sl;
        double thrq = 0.0;
        int doneq = 1;
        if (table) {
            sl = lSlot[qq];
            thrq = lThr[qq];
            doneq = lDone[qq];
        } else {
            const SlotParam& sp = a.sp[qq];
            sl = make_int4(sp.kind, sp.target, sp.wantParent, sp.gen);
        }
        const int kind = __builtin_amdgcn_readfirstlane(sl.x);
        const int t = __builtin_amdgcn_readfirstlane(sl.y);
        const int wantParent = __builtin_amdgcn_readfirstlane(sl.z);
        const int gen = __builtin_amdgcn_readfirstlane(sl.w);
        const bool chainSlot = chainMax > 0 && !doneq;
        const size_t sb = (size_t)qq * a.V;
        const size_t row = (size_t)vv * a.K;
        double ub = DBL_MAX;
        if (kind == Q_SSSD) ub = __longlong_as_double((long long)a.dist[sb + t]);
        else if (kind == Q_JOIN) ub = join_bound(a.sp[qq]);
        unsigned long long rowMin = ~0ull;        // smallest key left to a later round
        unsigned long long exitMin = ~0ull;       // JOIN: smallest exit sum this row made
        for (int b = 0; b < tl; b += 64 * kSteps) {
            int xs[kSteps];
            unsigned long long nb[kSteps], val[kSteps];
            // arcs: loads without branches (an index past the row re-reads
            // its last arc and is masked), so all of them are in flight at once
            double ws[kSteps];
            float hs[kSteps];
#pragma unroll
            for (int st = 0; st < kSteps; ++st) {
                const int i = min(b + st * 64 + lane, tl - 1);
                xs[st] = a.adj[row + i];
                ws[st] = a.w[row + i];
            }
            // target distances, and for A* the targets' bounds with them
#pragma unroll
            for (int st = 0; st < kSteps; ++st) {
                nb[st] = (unsigned long long)__double_as_longlong(dvv + ws[st]);
                val[st] = a.dist[sb + xs[st]];
            }
            if (kind == Q_SSSD || kind == Q_JOIN) {
#pragma unroll
                for (int st = 0; st < kSteps; ++st) hs[st] = a.hdist[sb + xs[st]];
            } else {
#pragma unroll
                for (int st = 0; st < kSteps; ++st) hs[st] = 0.0f;
            }
            // candidates (with parents kept, an exact tie is one too)
            unsigned cand = 0;
#pragma unroll
            for (int st = 0; st < kSteps; ++st)
                if (b + st * 64 + lane < tl && (wantParent ? nb[st] <= val[st] : nb[st] < val[st]))
                    cand |= 1u << st;
            // A*: drop a candidate that even the straight line to the target
            // cannot bring under the current answer (SVG.cu:399-405)
            if (kind == Q_SSSD) {
#pragma unroll
                for (int st = 0; st < kSteps; ++st)
                    if ((cand >> st) & 1u) {
                        const double nd = __longlong_as_double((long long)nb[st]);
                        if ((nd + (double)hs[st] * (1.0 - 1e-12)) > ub) cand &= ~(1u << st);
                    }
            } else if (kind == Q_JOIN) {
                // JOIN: nothing farther than the best exit sum can matter
#pragma unroll
                for (int st = 0; st < kSteps; ++st)
                    if ((cand >> st) & 1u) {
                        const double nd = __longlong_as_double((long long)nb[st]);
                        if (nd * (1.0 - 1e-12) > ub) cand &= ~(1u << st);
                    }
            }
            if (!__ballot(cand != 0)) continue;
            // atomicMins; val becomes the previous distance
#pragma unroll
            for (int st = 0; st < kSteps; ++st)
                if ((cand >> st) & 1u) val[st] = atomicMin(&a.dist[sb + xs[st]], nb[st]);
            unsigned imp = 0, win = 0;
#pragma unroll
            for (int st = 0; st < kSteps; ++st)
                if ((cand >> st) & 1u) {
                    if (nb[st] < val[st]) imp |= 1u << st;
                    else if (nb[st] == val[st] && wantParent) win |= 1u << st;
                }
            if (wantParent) win |= imp;
            if (kind == Q_JOIN) {
                // an improved exit lowers the bound: distance + offset (the
                // offset rounded up, so the bound stays at or above the exact sum)
#pragma unroll
                for (int st = 0; st < kSteps; ++st)
                    if (((imp >> st) & 1u) && hs[st] < __int_as_float(0x7f800000)) {
                        const unsigned long long xb = (unsigned long long)__double_as_longlong(
                            __longlong_as_double((long long)nb[st]) + (double)hs[st]);
                        exitMin = xb < exitMin ? xb : exitMin;
                    }
            }
            // val becomes the key (A*: distance + bound)
#pragma unroll
            for (int st = 0; st < kSteps; ++st)
                val[st] = kind == Q_SSSD ? (unsigned long long)__double_as_longlong(
                                               __longlong_as_double((long long)nb[st]) + (double)hs[st])
                                         : nb[st];
            // frontier flags of the improved targets
            unsigned own = 0;
#pragma unroll
            for (int st = 0; st < kSteps; ++st)
                if ((imp >> st) & 1u) {
                    const size_t o = sb + xs[st];
                    if (wantParent) {
                        a.stamp[o] = a.round;
                        a.parent[o] = INT_MAX;
                    }
                    // relaxed: nothing in this pass orders other data by the
                    // flag (an acquire/release here would write back and
                    // invalidate the L2 at every exchange)
                    if (__hip_atomic_exchange(&a.inq[o], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u)
                        own |= 1u << st;
                }
            // a target whose flag this wave took is chained (same bucket,
            // budget left) or pushed through the wave's LDS buffer; the keys
            // of improvements left to a later round feed the bucket look-ahead
#pragma unroll
            for (int st = 0; st < kSteps; ++st) {
                const bool im = (imp >> st) & 1u;
                const unsigned long long key = val[st];
                bool pu = (own >> st) & 1u;
                if (chainSlot) {
                    const bool ch = pu && __longlong_as_double((long long)key) < thrq;
                    const unsigned long long cm = __ballot(ch);
                    if (cm) {
                        const int room = chainMax - nChain;
                        const int rank = __popcll(cm & lanes_below(lane));
                        if (ch && rank < room) {
                            cq[wid][nChain + rank] = make_int2(qq, xs[st]);
                            pu = false;
                        }
                        nChain += min(room, __popcll(cm));
                        __builtin_amdgcn_wave_barrier();
                    }
                }
                if (im && (pu || !((own >> st) & 1u))) rowMin = key < rowMin ? key : rowMin;
                const unsigned long long pm = __ballot(pu);
                if (pm) {
                    if (nBuf + 64 > kPBuf) flush_pushes(a, buf, nBuf, list, lane);
                    if (pu) buf[nBuf + __popcll(pm & lanes_below(lane))] = pack_entry(gen, qq, xs[st]);
                    nBuf += __popcll(pm);
                    __builtin_amdgcn_wave_barrier();
                }
                const bool wi = (win >> st) & 1u;
                const unsigned long long wm = __ballot(wi);
                if (wm) {
                    const int wpos = won_pos(a, wonL, wonC, wBase, wUsed, wm, lane);
                    if (wi && (size_t)wpos < a.wonCap) {
                        WonEnt we;
                        we.slot = qq; we.x = xs[st]; we.u = vv; we.pad = 0; we.nb = nb[st];
                        wonL[wpos] = we;
                    }
                }
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const unsigned long long o = __shfl_xor(rowMin, off);
            rowMin = o < rowMin ? o : rowMin;
        }
        if (lane == 0 && rowMin != ~0ull) atomicMin(&acc[qq].minImp, rowMin);
        if (kind == Q_JOIN) {
            for (int off = 32; off > 0; off >>= 1) {
                const unsigned long long o = __shfl_xor(exitMin, off);
                exitMin = o < exitMin ? o : exitMin;
            }
            if (lane == 0 && exitMin != ~0ull) atomicMin((unsigned long long*)&a.sp[qq].result, exitMin);
        }
    };

    // the wave's rows, then the rows it chained (the queue grows while it is
    // worked through)
    int g = wave, h = 0;
    for (;;) {
        int qq, vv, tl;
        double dvv;
        if (g < n) {
            // the row's parameters are wave-uniform: keep them in scalar registers
            const int sgi = seg_of(pre, kCBlocks, g);
            const ExpEnt ex = a.exp[(size_t)sgi * stride + (g - pre[sgi])];
            g += nw;
            const uint64_t e = ex.entry;
            qq = __builtin_amdgcn_readfirstlane((int)((e >> 32) & 0xFFFFu));
            vv = __builtin_amdgcn_readfirstlane((int)(e & 0xFFFFFFFFu));
            tl = __builtin_amdgcn_readfirstlane((int)(e >> 48));
            dvv = __longlong_as_double((long long)ex.dist);
        } else if (h < nChain) {
            const int2 ce = cq[wid][h++];
            qq = __builtin_amdgcn_readfirstlane(ce.x);
            vv = __builtin_amdgcn_readfirstlane(ce.y);
            const size_t o = (size_t)qq * a.V + vv;
            unsigned long long d = 0;
            int t0 = 0;
            if (lane == 0) {
                // the flag is cleared before the distance is read (the read
                // waits on the exchange's result): an improvement after the
                // clear pushes the vertex again, one before it is in the value
                if (__hip_atomic_exchange(&a.inq[o], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <= 1u)
                    d = atomicMin(&a.dist[o], kInfBits);
                t0 = a.tail[vv];
                atomicOr(&acc[qq].expanded, 1);
                atomicAdd(&acc[qq].settled, 1ull);
                atomicAdd(&acc[qq].relaxed, (unsigned long long)t0);
            }
            dvv = __longlong_as_double((long long)__shfl(d, 0));
            tl = __builtin_amdgcn_readfirstlane(__shfl(t0, 0));
        } else {
            break;
        }
        relax_row(qq, vv, dvv, tl);
    }
    flush_pushes(a, buf, nBuf, list, lane);
    for (int j = wUsed + lane; j < kChunk; j += 64)
        if ((size_t)(wBase + j) < a.wonCap) wonL[wBase + j].slot = -1;
}

// Parents of this round's improvements, as a pass of its own: after the last
// round of a batch, and every round when the slots are bookkept by a pass.
__global__ void __launch_bounds__(kSolveBlock) k_parent_fix(SolveArgs a)
{
    __shared__ int pre[kCBlocks + 1];
    parent_part(a, a.c, a.round, a.won, a.exp, pre, blockIdx.x, gridDim.x);
}

// Bookkeeping of this round as a pass of its own (one block): after the last
// round of a batch, so the host sees current states, and every round when
// there are more slots than the classify pass bookkeeps in LDS.  It records
// the next round's bounds and clears the consumed statistics.
__global__ void k_bookkeep(SolveArgs a)
{
    for (int q = threadIdx.x; q < a.S; q += blockDim.x) {
        double t; int d;
        book_slot(a, q, a.acc, a.thr[a.c * a.S + q], a.doneP[a.c * a.S + q], t, d, true);
        a.thr[(a.c ^ 1) * a.S + q] = t;
        a.doneP[(a.c ^ 1) * a.S + q] = d;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kAccCopies * a.S; i += blockDim.x) {
        SlotAcc& x = a.acc[i];
        x.minDeferred = ~0ull; x.minImp = ~0ull; x.settled = 0; x.relaxed = 0; x.expanded = 0;
    }
}

__global__ void k_resolve_src(const int32_t* parent, int32_t* srcid, const unsigned long long* dist, int V,
                              int slot)
{
    const size_t sb = (size_t)slot * V;
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        if (dist[sb + v] == kInfBits) { srcid[sb + v] = -1; continue; }
        int u = v, hops = 0;
        while (hops++ <= V) {
            const int p = parent[sb + u];
            if (p == -1) break;
            if (p < 0 || p >= V) { u = -1; break; }
            u = p;
        }
        srcid[sb + v] = (hops > V) ? -1 : u;
    }
}

__global__ void k_mean_arc(const int32_t* tail, const double* w, int V, int K, double* sum, unsigned long long* cnt)
{
    double s = 0.0;
    unsigned long long c = 0;
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        const int t = tail[v];
        for (int i = 0; i < t; ++i) s += w[(size_t)v * K + i];
        c += (unsigned long long)t;
    }
    atomicAdd(sum, s);
    atomicAdd(cnt, c);
}

double graph_mean_arc(const DevGraph& g)
{
    // the bucket width only schedules work (answers do not depend on it):
    // a failed step falls back to width 1
    double* ds = nullptr;
    unsigned long long* dc = nullptr;
    double s = 0.0;
    unsigned long long c = 0;
    bool ok = hipMalloc(&ds, sizeof(double)) == hipSuccess && hipMalloc(&dc, sizeof(unsigned long long)) == hipSuccess &&
              hipMemset(ds, 0, sizeof(double)) == hipSuccess && hipMemset(dc, 0, sizeof(unsigned long long)) == hipSuccess;
    if (ok) {
        hipLaunchKernelGGL(k_mean_arc, dim3(256), dim3(256), 0, 0, g.tail, g.w, g.V, g.K, ds, dc);
        ok = hipMemcpy(&s, ds, sizeof(double), hipMemcpyDeviceToHost) == hipSuccess &&
             hipMemcpy(&c, dc, sizeof(unsigned long long), hipMemcpyDeviceToHost) == hipSuccess;
    }
    hipFree(ds); hipFree(dc);
    return ok && c ? s / (double)c : 1.0;
}

static SolveArgs make_args(SolverState& s, const DevGraph& g, const DevMesh& m)
{
    SolveArgs a;
    a.V = s.V; a.K = s.K;
    a.tail = g.tail; a.adj = g.adj; a.w = g.w; a.pos = m.pos;
    a.dist = s.dist; a.parent = s.parent; a.stamp = s.stamp; a.inq = s.inq; a.srcid = s.srcid;
    a.hdist = s.hdist;
    const int c = s.cur;
    a.c = c;
    a.dIn = s.dl[c]; a.dOut = s.dl[c ^ 1];
    a.pIn = s.pl[c]; a.pOut = s.pl[c ^ 1];
    a.exp = s.exp[c]; a.expPrev = s.exp[c ^ 1]; a.cnt = s.cnt;
    a.frontCap = s.frontCap; a.pCap = s.pCap;
    a.sp = s.sp; a.acc = s.acc[c]; a.accPrev = s.acc[c ^ 1]; a.S = s.S; a.round = s.round;
    a.thr = s.thr; a.doneP = s.doneP;
    a.won = s.won[c]; a.wonPrev = s.won[c ^ 1]; a.wonCap = s.wonCap;
    a.doBook = 0; a.doParent = 0;
    a.ovFlag = s.ovDev;
    a.chainBudget = s.chainBudget;
    return a;
}

#define HTRY(x)                                                                   \
    do {                                                                          \
        hipError_t _e = (x);                                                      \
        if (_e != hipSuccess) return std::string(#x ": ") + hipGetErrorString(_e); \
    } while (0)

size_t solver_bytes_per_slot(int V)
{
    // dist, parent, stamp, inq, srcid, hdist; deferred lists (2) and row
    // lists (2, 16 B) at two entries per vertex; push lists (2 x kPLists) at
    // one; won lists (2 x kPLists) at a quarter
    return (size_t)V * (8 + 4 + 4 + 4 + 4 + 4 + 2 * 2 * 8 + 2 * 2 * 16 + 2 * kPLists * 8 +
                        2 * kPLists * sizeof(WonEnt) / 4);
}

std::string solver_create(const DevGraph& g, const DevMesh& m, int slots, double delta, SolverState& s)
{
    (void)m;
    if (slots < 1 || slots > 4096) return "slots must be in [1, 4096]";
    if (g.K > 0xFFFF) return "K above 65535: a row's length must fit 16 bits of its list entry";
    s.V = g.V; s.K = g.K; s.S = slots; s.delta = delta; s.deltaMsad = delta;
    const size_t sv = (size_t)slots * (size_t)g.V;
    // live entries are at most one per (slot, vertex); entries of a slot's
    // previous query may still be listed once more
    s.frontCap = 2 * sv + 65536;
    s.pCap = sv + 65536;
    if (s.frontCap >= 0x7fffffffull) return "slots * V too large for the frontier";
    HTRY(hipMalloc(&s.dist, sizeof(unsigned long long) * sv));
    HTRY(hipMalloc(&s.parent, sizeof(int32_t) * sv));
    HTRY(hipMalloc(&s.stamp, sizeof(uint32_t) * sv));
    HTRY(hipMalloc(&s.inq, sizeof(uint32_t) * sv));
    HTRY(hipMalloc(&s.srcid, sizeof(int32_t) * sv));
    HTRY(hipMalloc(&s.hdist, sizeof(float) * sv));
    for (int i = 0; i < 2; ++i) {
        HTRY(hipMalloc(&s.dl[i], sizeof(uint64_t) * s.frontCap));
        HTRY(hipMalloc(&s.pl[i], sizeof(uint64_t) * s.pCap * kPLists));
    }
    for (int i = 0; i < 2; ++i) HTRY(hipMalloc(&s.exp[i], sizeof(ExpEnt) * s.frontCap));
    HTRY(hipMalloc(&s.cnt, sizeof(SolveCounts)));
    HTRY(hipMemset(s.cnt, 0, sizeof(SolveCounts)));
    s.wonCap = std::min<size_t>(sv / 4 + 65536, 0x7ffffff0ull);
    for (int i = 0; i < 2; ++i) HTRY(hipMalloc(&s.won[i], sizeof(WonEnt) * s.wonCap * kPLists));
    {
        std::vector<SlotAcc> a0((size_t)kAccCopies * slots);
        for (auto& x : a0) { x.minDeferred = ~0ull; x.minImp = ~0ull; x.settled = 0; x.relaxed = 0; x.expanded = 0; x.pad = 0; }
        for (int i = 0; i < 2; ++i) {
            HTRY(hipMalloc(&s.acc[i], sizeof(SlotAcc) * a0.size()));
            HTRY(hipMemcpy(s.acc[i], a0.data(), sizeof(SlotAcc) * a0.size(), hipMemcpyHostToDevice));
        }
    }
    HTRY(hipMalloc(&s.thr, sizeof(double) * 2 * (size_t)slots));
    HTRY(hipMemset(s.thr, 0, sizeof(double) * 2 * (size_t)slots));
    HTRY(hipMalloc(&s.doneP, sizeof(int) * 2 * (size_t)slots));
    {
        std::vector<int> d1((size_t)2 * slots, 1);
        HTRY(hipMemcpy(s.doneP, d1.data(), sizeof(int) * d1.size(), hipMemcpyHostToDevice));
    }
    HTRY(hipMalloc(&s.sp, sizeof(SlotParam) * (size_t)(slots + 1)));
    HTRY(hipHostMalloc(&s.spHost, sizeof(SlotParam) * (size_t)(slots + 1), hipHostMallocDefault));
    std::vector<SlotParam> init((size_t)slots + 1);
    for (auto& p : init) { std::memset(&p, 0, sizeof(p)); p.kind = -1; p.done = 1; p.minDeferred = ~0ull; }
    init[slots].kind = 0;                           // the overflow flag
    HTRY(hipMemcpy(s.sp, init.data(), sizeof(SlotParam) * init.size(), hipMemcpyHostToDevice));
    std::memcpy(s.spHost, init.data(), sizeof(SlotParam) * init.size());
    s.ovDev = &s.sp[slots].kind;
    for (int i = 0; i < 2; ++i) HTRY(hipEventCreateWithFlags(&s.stageEv[i], hipEventDisableTiming));
    s.cur = 0;
    s.round = 1;
    s.gens.assign((size_t)slots, 0);
    // the set-up copies and fills above ran on the null stream; s.stream may
    // be a non-blocking stream that does not wait for them
    HTRY(hipDeviceSynchronize());
    return std::string();
}

void solver_destroy(SolverState& s)
{
    hipFree(s.dist); hipFree(s.parent); hipFree(s.stamp); hipFree(s.inq); hipFree(s.srcid); hipFree(s.hdist);
    for (int i = 0; i < 2; ++i) { hipFree(s.dl[i]); hipFree(s.pl[i]); }
    for (int i = 0; i < 2; ++i) { hipFree(s.exp[i]); hipFree(s.won[i]); hipFree(s.acc[i]); }
    hipFree(s.cnt); hipFree(s.thr); hipFree(s.doneP);
    hipFree(s.sp); hipHostFree(s.spHost);
    for (int i = 0; i < 2; ++i) {
        hipFree(s.stage[i]); hipHostFree(s.stageHost[i]);
        if (s.stageEv[i]) hipEventDestroy(s.stageEv[i]);
    }
    s = SolverState();
}

// Seeds of the admitted slots, one block each, in order (SVG.cu:246-249,
// 359-365): a repeated source keeps its last weight, as the reference's
// array writes do.  Seeds join push list 0 of the next round's input.  The
// first bucket bound is the smallest seed key + delta.
__global__ void k_admit(SolveArgs a, const int* __restrict__ slots, const int* __restrict__ kinds,
                        const int* __restrict__ targets, const int* __restrict__ wantP,
                        const int* __restrict__ gens, const int* __restrict__ off,
                        const int* __restrict__ seeds, const double* __restrict__ wts, double deltaSssd,
                        double deltaMsad, const int* __restrict__ xoff, const int* __restrict__ xv,
                        const double* __restrict__ xw, float* hdistW)
{
    const int k = blockIdx.x;
    const int slot = slots[k], kind = kinds[k], target = targets[k], gen = gens[k];
    const size_t base = (size_t)slot * a.V;
    const double delta = kind == Q_SSSD ? deltaSssd : deltaMsad;
    uint64_t* front = const_cast<uint64_t*>(a.pIn);
    double mk = DBL_MAX, xb = __int_as_float(0x7f800000);
    if (kind == Q_JOIN) {
        // exit offsets, rounded up so that sums with them bound the exact ones
        for (int i = xoff[k] + threadIdx.x; i < xoff[k + 1]; i += blockDim.x)
            hdistW[base + xv[i]] = __double2float_ru(xw[i]);
        __syncthreads();
        // the seeds are distinct (one run's vertex table): one per lane of the
        // block's single wave, one list position reservation per 64
        const int lane = threadIdx.x;
        for (int i0 = off[k]; i0 < off[k + 1]; i0 += 64) {
            const int i = i0 + lane;
            const bool act = i < off[k + 1];
            int v = 0;
            if (act) {
                v = seeds[i];
                const double w = wts[i];
                a.dist[base + v] = (unsigned long long)__double_as_longlong(w);
                a.parent[base + v] = -1;
                a.srcid[base + v] = v;
                a.inq[base + v] = 1;
                mk = w < mk ? w : mk;
                const double x = w + (double)hdistW[base + v];
                xb = x < xb ? x : xb;
            }
            const unsigned long long m = __ballot(act);
            int pos0 = 0;
            if (lane == 0) pos0 = atomicAdd(&a.cnt->pcnt[a.c][0], __popcll(m));
            pos0 = __shfl(pos0, 0);
            if (act) {
                const int pos = pos0 + __popcll(m & lanes_below(lane));
                if ((size_t)pos < a.pCap) front[pos] = pack_entry(gen, slot, v);
                else atomicOr(a.ovFlag, 1);
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            const double om = __shfl_xor(mk, o), ox = __shfl_xor(xb, o);
            mk = om < mk ? om : mk;
            xb = ox < xb ? ox : xb;
        }
    }
    if (threadIdx.x != 0) return;
    for (int i = off[k]; kind != Q_JOIN && i < off[k + 1]; ++i) {
        const int v = seeds[i];
        const double w = wts[i];
        a.dist[base + v] = (unsigned long long)__double_as_longlong(w);
        a.parent[base + v] = -1;
        a.srcid[base + v] = v;
        if (a.inq[base + v] == 0) {
            a.inq[base + v] = 1;
            const int pos = atomicAdd(&a.cnt->pcnt[a.c][0], 1);
            if ((size_t)pos < a.pCap) front[pos] = pack_entry(gen, slot, v);
            else atomicOr(a.ovFlag, 1);
        }
        const double key = kind == Q_SSSD ? w + heur(a.pos, v, target) : w;
        mk = key < mk ? key : mk;
    }
    SlotParam p;
    p.kind = kind; p.target = target; p.wantParent = wantP[k]; p.done = 0;
    p.thr = mk + delta; p.delta = delta; p.minDeferred = ~0ull;
    p.expanded = 0; p.rounds = 0; p.gen = gen; p.pad = 0; p.relaxed = 0; p.settled = 0;
    p.result = kind == Q_JOIN ? xb : DBL_MAX;
    a.sp[slot] = p;
    a.thr[a.c * a.S + slot] = p.thr;
    a.doneP[a.c * a.S + slot] = 0;
}

std::string solver_admit_many(SolverState& s, const DevMesh& m, const std::vector<Admission>& adm)
{
    const int n = (int)adm.size();
    if (n == 0) return std::string();
    // the staging pair used two admissions ago is free once its copy is done
    const int si = s.stageIdx;
    s.stageIdx ^= 1;
    HTRY(hipEventSynchronize(s.stageEv[si]));
    // slots, kinds, targets, wantP, gens, off[n+1], seeds, xoff[n+1], exits
    std::vector<int> ints;
    ints.reserve((size_t)7 * n + 64);
    size_t nseeds = 0, nexits = 0;
    for (const Admission& a : adm) {
        if (a.slot < 0 || a.slot >= s.S) return "slot out of range";
        if (a.src.empty() || a.src.size() != a.w.size()) return "bad seeds";
        for (int v : a.src) if (v < 0 || v >= s.V) return "source out of range";
        if (a.kind == Q_SSSD && (a.target < 0 || a.target >= s.V)) return "target out of range";
        if (a.exitV.size() != a.exitOff.size()) return "bad exits";
        for (int v : a.exitV) if (v < 0 || v >= s.V) return "exit out of range";
        nseeds += a.src.size();
        nexits += a.kind == Q_JOIN ? a.exitV.size() : 0;
    }
    for (const Admission& a : adm) ints.push_back(a.slot);
    for (const Admission& a : adm) ints.push_back(a.kind);
    for (const Admission& a : adm) ints.push_back(a.kind == Q_SSSD ? a.target : (a.kind == Q_JOIN ? -2 : -1));
    for (const Admission& a : adm) ints.push_back(a.wantParent ? 1 : 0);
    for (const Admission& a : adm) {
        s.gens[a.slot] = (s.gens[a.slot] + 1) & 0xFFFF;
        ints.push_back(s.gens[a.slot]);
    }
    int o = 0;
    for (const Admission& a : adm) { ints.push_back(o); o += (int)a.src.size(); }
    ints.push_back(o);
    for (const Admission& a : adm) ints.insert(ints.end(), a.src.begin(), a.src.end());
    const size_t xoffPos = ints.size();
    o = 0;
    for (const Admission& a : adm) { ints.push_back(o); o += a.kind == Q_JOIN ? (int)a.exitV.size() : 0; }
    ints.push_back(o);
    for (const Admission& a : adm)
        if (a.kind == Q_JOIN) ints.insert(ints.end(), a.exitV.begin(), a.exitV.end());
    const size_t intBytes = ((ints.size() * sizeof(int) + 15) / 16) * 16;
    const size_t bytes = intBytes + (nseeds + nexits) * sizeof(double);
    if (bytes > s.stageCap[si]) {
        HTRY(hipStreamSynchronize(s.stream));       // the old buffer may still be read
        hipFree(s.stage[si]);
        s.stageCap[si] = std::max(bytes, (size_t)1 << 20);
        HTRY(hipMalloc(&s.stage[si], s.stageCap[si]));
        hipHostFree(s.stageHost[si]);
        HTRY(hipHostMalloc(&s.stageHost[si], s.stageCap[si], hipHostMallocDefault));
    }
    std::memcpy(s.stageHost[si], ints.data(), ints.size() * sizeof(int));
    double* wh = (double*)((char*)s.stageHost[si] + intBytes);
    for (const Admission& a : adm) for (double w : a.w) *wh++ = w;
    for (const Admission& a : adm)
        if (a.kind == Q_JOIN) for (double w : a.exitOff) *wh++ = w;
    HTRY(hipMemcpyAsync(s.stage[si], s.stageHost[si], bytes, hipMemcpyHostToDevice, s.stream));
    HTRY(hipEventRecord(s.stageEv[si], s.stream));
    const int* d = (const int*)s.stage[si];
    const int* dSlots = d;
    const int* dKinds = d + n;
    const int* dTargets = d + 2 * n;
    const int* dWant = d + 3 * n;
    const int* dGens = d + 4 * n;
    const int* dOff = d + 5 * n;
    const int* dSeeds = d + 6 * n + 1;
    const double* dW = (const double*)((const char*)s.stage[si] + intBytes);
    const int* dXoff = d + xoffPos;
    const int* dXv = d + xoffPos + n + 1;
    const double* dXw = dW + nseeds;
    const size_t total = (size_t)n * s.V;
    hipLaunchKernelGGL(k_slot_init, dim3((unsigned)std::min<size_t>(2048, (total + 255) / 256)), dim3(256), 0, s.stream,
                       s.dist, s.parent, s.stamp, s.inq, s.srcid, s.hdist, m.pos, dSlots, dTargets, n, s.V);
    DevGraph dummy;
    SolveArgs args = make_args(s, dummy, m);
    hipLaunchKernelGGL(k_admit, dim3(n), dim3(64), 0, s.stream, args, dSlots, dKinds, dTargets, dWant, dGens, dOff,
                       dSeeds, dW, s.delta, s.deltaMsad, dXoff, dXv, dXw, s.hdist);
    s.launches += 2;
    for (const Admission& a : adm) {
        SlotParam& p = s.spHost[a.slot];
        p.kind = a.kind; p.target = a.target; p.done = 0; p.wantParent = a.wantParent ? 1 : 0;
        p.gen = s.gens[a.slot];
    }
    return std::string();
}

std::string solver_admit(SolverState& s, const DevMesh& m, int slot, int kind, const int* src,
                         const double* w, int ns, int target, bool wantParent)
{
    if (ns <= 0) return "no sources";
    std::vector<Admission> adm(1);
    adm[0].slot = slot; adm[0].kind = kind; adm[0].target = target; adm[0].wantParent = wantParent;
    adm[0].src.assign(src, src + ns);
    adm[0].w.assign(w, w + ns);
    return solver_admit_many(s, m, adm);
}

std::string solver_enqueue_rounds(SolverState& s, const DevGraph& g, const DevMesh& m, int nRounds)
{
    s.anyParent = false;
    for (int q = 0; q < s.S; ++q)
        if (s.spHost[q].kind >= 0 && !s.spHost[q].done && s.spHost[q].wantParent) s.anyParent = true;
    if (nRounds < 1) nRounds = 1;
    // With few slots, a round is two launches: classify (which also does the
    // previous round's bookkeeping and parents) and expand; the batch ends
    // with the last round's parents and bookkeeping as passes of their own.
    const bool inClassify = s.S <= kLdsSlots;
    for (int r = 0; r < nRounds; ++r) {
        SolveArgs a = make_args(s, g, m);
        a.doBook = inClassify && r > 0;
        a.doParent = a.doBook && s.anyParent;
        hipLaunchKernelGGL(k_classify, dim3(kCBlocks), dim3(kSolveBlock), 0, s.stream, a);
        hipLaunchKernelGGL(k_expand, dim3(s.expandBlocks), dim3(kSolveBlock), 0, s.stream, a);
        s.launches += 2;
        if (!inClassify || r == nRounds - 1) {
            if (s.anyParent) {
                hipLaunchKernelGGL(k_parent_fix, dim3(1024), dim3(kSolveBlock), 0, s.stream, a);
                s.launches++;
            }
            hipLaunchKernelGGL(k_bookkeep, dim3(1), dim3(1024), 0, s.stream, a);
            s.launches++;
        }
        s.cur ^= 1;
        s.round++;
    }
    HTRY(hipGetLastError());
    return std::string();
}

std::string solver_enqueue_status(SolverState& s, SlotParam* hostDst)
{
    HTRY(hipMemcpyAsync(hostDst, s.sp, sizeof(SlotParam) * (size_t)(s.S + 1), hipMemcpyDeviceToHost, s.stream));
    return std::string();
}

std::string solver_rounds(SolverState& s, const DevGraph& g, const DevMesh& m, int nRounds, int& running)
{
    std::string e = solver_enqueue_rounds(s, g, m, nRounds);
    if (e.empty()) e = solver_enqueue_status(s, s.spHost);
    if (!e.empty()) return e;
    const hipError_t he = hipStreamSynchronize(s.stream);
    if (he != hipSuccess) return std::string("solver round: ") + hipGetErrorString(he);
    if (s.spHost[s.S].kind) return "solver frontier overflow";
    running = 0;
    for (int q = 0; q < s.S; ++q)
        if (s.spHost[q].kind >= 0 && !s.spHost[q].done) ++running;
    return std::string();
}

std::string solver_enqueue_fetch(SolverState& s, int slot, hipStream_t stream, double* distDst,
                                 int32_t* srcidDst, int32_t* parentDst)
{
    const size_t sb = (size_t)slot * s.V;
    if (distDst)
        HTRY(hipMemcpyAsync(distDst, s.dist + sb, sizeof(double) * (size_t)s.V, hipMemcpyDeviceToHost, stream));
    if (srcidDst) {
        hipLaunchKernelGGL(k_resolve_src, dim3((s.V + 255) / 256), dim3(256), 0, stream, s.parent, s.srcid, s.dist,
                           s.V, slot);
        HTRY(hipMemcpyAsync(srcidDst, s.srcid + sb, sizeof(int32_t) * (size_t)s.V, hipMemcpyDeviceToHost, stream));
    }
    if (parentDst)
        HTRY(hipMemcpyAsync(parentDst, s.parent + sb, sizeof(int32_t) * (size_t)s.V, hipMemcpyDeviceToHost, stream));
    return std::string();
}

std::string solver_fetch(SolverState& s, int slot, double* distOut, double* distT, int32_t* srcidOut,
                         int32_t* parentOut)
{
    const size_t sb = (size_t)slot * s.V;
    HTRY(hipStreamSynchronize(s.stream));
    if (distT) {
        unsigned long long b;
        HTRY(hipMemcpy(&b, s.dist + sb + s.spHost[slot].target, sizeof(b), hipMemcpyDeviceToHost));
        std::memcpy(distT, &b, sizeof(double));
    }
    if (distOut) HTRY(hipMemcpy(distOut, s.dist + sb, sizeof(double) * (size_t)s.V, hipMemcpyDeviceToHost));
    if (srcidOut) {
        hipLaunchKernelGGL(k_resolve_src, dim3((s.V + 255) / 256), dim3(256), 0, s.stream, s.parent, s.srcid, s.dist,
                           s.V, slot);
        HTRY(hipStreamSynchronize(s.stream));
        HTRY(hipMemcpy(srcidOut, s.srcid + sb, sizeof(int32_t) * (size_t)s.V, hipMemcpyDeviceToHost));
    }
    if (parentOut) HTRY(hipMemcpy(parentOut, s.parent + sb, sizeof(int32_t) * (size_t)s.V, hipMemcpyDeviceToHost));
    return std::string();
}

// A finished slot stays "done" on the device (its entries are dropped) until
// it is admitted again; releasing it is host bookkeeping only.
void solver_release_slot(SolverState& s, int slot)
{
    s.spHost[slot].kind = -1;
}

}  // namespace svg

using namespace svg;

void svg_solver_release(svg_graph_s* g)
{
    if (g && g->solver) {
        solver_destroy(*g->solver);
        delete g->solver;
        g->solver = nullptr;
    }
}

void svg_point_release(svg_graph_s* g)
{
    if (g && g->point) {
        point_free(*g->point);
        delete g->point;
        g->point = nullptr;
    }
}

static int get_solver(svg_graph_s* g, int slots, SolverState** out)
{
    int rc = svg_ensure_device();
    if (rc) return rc;
    if (g->solver && g->solver->S < slots) svg_solver_release(g);
    if (!g->solver) {
        g->solver = new SolverState();
        const double delta = graph_mean_arc(g->d);
        std::string e = solver_create(g->d, g->mesh->d, slots, delta, *g->solver);
        if (!e.empty()) { svg_solver_release(g); return svg_set_error(SVG_E_MEM, e); }
    }
    *out = g->solver;
    return SVG_OK;
}

static int run_to_completion(svg_graph_s* g, SolverState* s)
{
    for (;;) {
        int running = 0;
        std::string e = solver_round(*s, g->d, g->mesh->d, running);
        if (!e.empty()) return svg_set_error(SVG_E_HIP, e);
        if (running == 0) return SVG_OK;
    }
}

extern "C" int svg_sssd(svg_graph_t g, int s, int t, double* dist_t, int32_t* path, int path_cap, int* path_len)
{
    return svg_guarded<int>([&]() -> int {
        if (!g || !dist_t) return svg_set_error(SVG_E_ARG, "null argument");
        if (s < 0 || t < 0 || s >= g->d.V || t >= g->d.V) return svg_set_error(SVG_E_ARG, "vertex out of range");
        std::lock_guard<std::mutex> lk(g->mu);
        SolverState* st = nullptr;
        int rc = get_solver(g, 1, &st);
        if (rc) return rc;
        const double w0 = 0.0;
        const bool wantPath = path && path_cap > 0;
        std::string e = solver_admit(*st, g->mesh->d, 0, Q_SSSD, &s, &w0, 1, t, wantPath);
        if (!e.empty()) return svg_set_error(SVG_E_ARG, e);
        rc = run_to_completion(g, st);
        if (rc) return rc;
        std::vector<int32_t> par;
        if (wantPath) par.resize((size_t)g->d.V);
        e = solver_fetch(*st, 0, nullptr, dist_t, nullptr, wantPath ? par.data() : nullptr);
        solver_release_slot(*st, 0);
        if (!e.empty()) return svg_set_error(SVG_E_HIP, e);
        if (wantPath) {
            int n = 0, u = t;
            if (*dist_t < DBL_MAX) {
                while (u != -1 && n < path_cap) {
                    path[n++] = u;
                    if (u == s) break;
                    u = par[u];
                    if (u == INT_MAX) { u = -1; }
                }
            }
            if (path_len) *path_len = n;
        }
        return SVG_OK;
    });
}

extern "C" int svg_msad(svg_graph_t g, const int32_t* src, const double* w, int ns, double* dist, int32_t* src_id,
                        int32_t* parent)
{
    return svg_guarded<int>([&]() -> int {
        if (!g || !src || !w || !dist || ns <= 0) return svg_set_error(SVG_E_ARG, "bad argument");
        for (int i = 0; i < ns; ++i)
            if (!std::isfinite(w[i]) || std::signbit(w[i])) return svg_set_error(SVG_E_ARG, kBadWeight);
        std::lock_guard<std::mutex> lk(g->mu);
        SolverState* st = nullptr;
        int rc = get_solver(g, 1, &st);
        if (rc) return rc;
        const bool wantParent = src_id || parent;
        std::string e = solver_admit(*st, g->mesh->d, 0, Q_MSAD, src, w, ns, -1, wantParent);
        if (!e.empty()) return svg_set_error(SVG_E_ARG, e);
        rc = run_to_completion(g, st);
        if (rc) return rc;
        e = solver_fetch(*st, 0, dist, nullptr, src_id, parent);
        solver_release_slot(*st, 0);
        if (!e.empty()) return svg_set_error(SVG_E_HIP, e);
        return SVG_OK;
    });
}

// ---------------------------------------------------------------------------
// MSAD without device allocation (svg_msad_noalloc).  Every device buffer of
// one single-slot solver -- the per-vertex state, the frontier lists, the
// round counters, the slot record, the admission staging and the mean-arc
// accumulators -- is carved out of `scratch`, a file-scope device array of
// kNoAllocBytes that each GPU holds for the life of the process.  The same
// kernels run with the same list capacities as a solver made by
// solver_create for one slot, so the answers are those of svg_msad.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocBytes = (size_t)64 << 20;
constexpr int kNoAllocDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

// Device address of `scratch` on each device, looked up on the first call
// made with that device current (the host-side name is not a device
// address).
void* g_scratchDev[kNoAllocDevices] = {nullptr};
std::mutex g_scratchMu;            // calls are serialised: they share the array

struct NoAllocLayout {
    size_t off[24];
    size_t total;
};

// Offsets (256-B aligned) of every buffer for V vertices and ns seeds; the
// order matches the fields assigned in noalloc_carve.
NoAllocLayout noalloc_layout(int V, int ns)
{
    NoAllocLayout L;
    const size_t sv = (size_t)V;
    const size_t frontCap = 2 * sv + 65536, pCap = sv + 65536;
    const size_t wonCap = std::min<size_t>(sv / 4 + 65536, 0x7ffffff0ull);
    const size_t intBytes = (((size_t)7 + (size_t)ns + 2) * sizeof(int) + 15) / 16 * 16;
    const size_t sizes[] = {
        sizeof(unsigned long long) * sv, sizeof(int32_t) * sv, sizeof(uint32_t) * sv, sizeof(uint32_t) * sv,
        sizeof(int32_t) * sv, sizeof(float) * sv,                                   // 0-5 per-vertex state
        sizeof(uint64_t) * frontCap, sizeof(uint64_t) * frontCap,                   // 6-7 deferred lists
        sizeof(uint64_t) * pCap * kPLists, sizeof(uint64_t) * pCap * kPLists,       // 8-9 push lists
        sizeof(ExpEnt) * frontCap, sizeof(ExpEnt) * frontCap,                       // 10-11 rows
        sizeof(SolveCounts),                                                        // 12
        sizeof(WonEnt) * wonCap * kPLists, sizeof(WonEnt) * wonCap * kPLists,       // 13-14 won lists
        sizeof(SlotAcc) * kAccCopies, sizeof(SlotAcc) * kAccCopies,                 // 15-16 statistics
        sizeof(double) * 2, sizeof(int) * 2,                                        // 17-18 thr, doneP
        sizeof(SlotParam) * 2,                                                      // 19 slot record + flag
        intBytes + sizeof(double) * (size_t)ns,                                     // 20 admission staging
        sizeof(double) + sizeof(unsigned long long),                                // 21 mean-arc sum, count
    };
    size_t o = 0;
    for (int i = 0; i < 22; ++i) {
        L.off[i] = o;
        o += (sizes[i] + 255) / 256 * 256;
    }
    L.total = o;
    return L;
}

}  // namespace

// The state a fresh single-slot solver starts from (solver_create's
// initial values), written by a kernel: the array keeps the last call's
// contents, and fills are kernels, not memsets.
__global__ void k_noalloc_reset(SolveCounts* cnt, SlotAcc* acc0, SlotAcc* acc1, double* thr, int* doneP,
                                SlotParam* sp, double* msum, unsigned long long* mcnt)
{
    int* c = reinterpret_cast<int*>(cnt);
    const int nc = (int)(sizeof(SolveCounts) / sizeof(int));
    for (int i = threadIdx.x; i < nc; i += blockDim.x) c[i] = 0;
    for (int i = threadIdx.x; i < kAccCopies; i += blockDim.x) {
        SlotAcc x;
        x.minDeferred = ~0ull; x.minImp = ~0ull; x.settled = 0; x.relaxed = 0; x.expanded = 0; x.pad = 0;
        acc0[i] = x;
        acc1[i] = x;
    }
    if (threadIdx.x < 2) { thr[threadIdx.x] = 0.0; doneP[threadIdx.x] = 1; }
    if (threadIdx.x < 2) {
        SlotParam p;
        p.kind = threadIdx.x == 0 ? -1 : 0;        // the second record is the overflow flag
        p.target = 0; p.wantParent = 0; p.done = 1; p.thr = 0.0; p.delta = 0.0; p.minDeferred = ~0ull;
        p.expanded = 0; p.rounds = 0; p.gen = 0; p.pad = 0; p.relaxed = 0; p.settled = 0; p.result = 0.0;
        sp[threadIdx.x] = p;
    }
    if (threadIdx.x == 0) { *msum = 0.0; *mcnt = 0ull; }
}

extern "C" int svg_msad_noalloc(svg_graph_t g, const int32_t* src, const double* w, int ns, double* dist,
                                int32_t* src_id, int32_t* parent)
{
    return svg_guarded<int>([&]() -> int {
        if (!g || !src || !w || !dist || ns <= 0) return svg_set_error(SVG_E_ARG, "bad argument");
        const int V = g->d.V;
        for (int i = 0; i < ns; ++i) {
            if (src[i] < 0 || src[i] >= V) return svg_set_error(SVG_E_ARG, "source out of range");
            if (!std::isfinite(w[i]) || std::signbit(w[i])) return svg_set_error(SVG_E_ARG, kBadWeight);
        }
        if (g->d.K > 0xFFFF) return svg_set_error(SVG_E_ARG, "K above 65535");
        const NoAllocLayout L = noalloc_layout(V, ns);
        if (L.total > kNoAllocBytes)
            return svg_set_error(SVG_E_ARG, "graph too large for svg_msad_noalloc: it needs " + std::to_string(L.total) +
                                                " bytes of its " + std::to_string(kNoAllocBytes) +
                                                "-byte device array (about 87,000 vertices at most)");
        int rc = svg_ensure_device();
        if (rc) return rc;
        std::lock_guard<std::mutex> lk(g_scratchMu);
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNoAllocDevices)
            return svg_set_error(SVG_E_HIP, "no current device");
        if (!g_scratchDev[dev]) {
            void* p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p)
                return svg_set_error(SVG_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            g_scratchDev[dev] = p;
        }
        char* base = static_cast<char*>(g_scratchDev[dev]);
        SolverState s;
        s.V = V; s.K = g->d.K; s.S = 1; s.stream = 0;
        s.dist = (unsigned long long*)(base + L.off[0]);
        s.parent = (int32_t*)(base + L.off[1]);
        s.stamp = (uint32_t*)(base + L.off[2]);
        s.inq = (uint32_t*)(base + L.off[3]);
        s.srcid = (int32_t*)(base + L.off[4]);
        s.hdist = (float*)(base + L.off[5]);
        s.dl[0] = (uint64_t*)(base + L.off[6]); s.dl[1] = (uint64_t*)(base + L.off[7]);
        s.pl[0] = (uint64_t*)(base + L.off[8]); s.pl[1] = (uint64_t*)(base + L.off[9]);
        s.exp[0] = (ExpEnt*)(base + L.off[10]); s.exp[1] = (ExpEnt*)(base + L.off[11]);
        s.cnt = (SolveCounts*)(base + L.off[12]);
        s.won[0] = (WonEnt*)(base + L.off[13]); s.won[1] = (WonEnt*)(base + L.off[14]);
        s.acc[0] = (SlotAcc*)(base + L.off[15]); s.acc[1] = (SlotAcc*)(base + L.off[16]);
        s.thr = (double*)(base + L.off[17]);
        s.doneP = (int*)(base + L.off[18]);
        s.sp = (SlotParam*)(base + L.off[19]);
        s.ovDev = &s.sp[1].kind;
        s.frontCap = 2 * (size_t)V + 65536;
        s.pCap = (size_t)V + 65536;
        s.wonCap = std::min<size_t>((size_t)V / 4 + 65536, 0x7ffffff0ull);
        int* stage = (int*)(base + L.off[20]);
        const size_t intBytes = (((size_t)7 + (size_t)ns + 2) * sizeof(int) + 15) / 16 * 16;
        double* stageW = (double*)((char*)stage + intBytes);
        double* msum = (double*)(base + L.off[21]);
        unsigned long long* mcnt = (unsigned long long*)(msum + 1);
        // host copy of the slot record and overflow flag (no pinned memory)
        SlotParam spHost[2];
        std::memset(spHost, 0, sizeof(spHost));
        spHost[0].kind = -1; spHost[0].done = 1; spHost[1].kind = 0;
        s.spHost = spHost;

        hipLaunchKernelGGL(k_noalloc_reset, dim3(1), dim3(256), 0, 0, s.cnt, s.acc[0], s.acc[1], s.thr, s.doneP, s.sp,
                           msum, mcnt);
        // bucket width: the graph's mean arc length, as graph_mean_arc computes it
        hipLaunchKernelGGL(k_mean_arc, dim3(256), dim3(256), 0, 0, g->d.tail, g->d.w, V, g->d.K, msum, mcnt);
        double msumH = 0.0;
        unsigned long long mcntH = 0;
        hipError_t he = hipMemcpy(&msumH, msum, sizeof(double), hipMemcpyDeviceToHost);
        if (he == hipSuccess) he = hipMemcpy(&mcntH, mcnt, sizeof(mcntH), hipMemcpyDeviceToHost);
        if (he != hipSuccess) return svg_set_error(SVG_E_HIP, std::string("svg_msad_noalloc set-up: ") + hipGetErrorString(he));
        s.delta = s.deltaMsad = mcntH ? msumH / (double)mcntH : 1.0;
        // admission staging, laid out as solver_admit_many lays it out for one
        // slot: slot, kind, target, wantParent, gen, off[2], seeds, xoff[2]; weights
        const int wantParent = (src_id || parent) ? 1 : 0;
        const int head[7] = {0, Q_MSAD, -1, wantParent, 1, 0, ns};
        const int tailOff[2] = {0, 0};
        he = hipMemcpy(stage, head, sizeof(head), hipMemcpyHostToDevice);
        if (he == hipSuccess) he = hipMemcpy(stage + 7, src, sizeof(int32_t) * (size_t)ns, hipMemcpyHostToDevice);
        if (he == hipSuccess) he = hipMemcpy(stage + 7 + ns, tailOff, sizeof(tailOff), hipMemcpyHostToDevice);
        if (he == hipSuccess) he = hipMemcpy(stageW, w, sizeof(double) * (size_t)ns, hipMemcpyHostToDevice);
        if (he != hipSuccess) return svg_set_error(SVG_E_HIP, std::string("svg_msad_noalloc seeds: ") + hipGetErrorString(he));
        const DevMesh& m = g->mesh->d;
        const int* dSlots = stage;
        const int* dKinds = stage + 1;
        const int* dTargets = stage + 2;
        const int* dWant = stage + 3;
        const int* dGens = stage + 4;
        const int* dOff = stage + 5;
        const int* dSeeds = stage + 7;
        const int* dXoff = stage + 7 + ns;
        hipLaunchKernelGGL(k_slot_init, dim3((unsigned)std::min<size_t>(2048, ((size_t)V + 255) / 256)), dim3(256), 0, 0,
                           s.dist, s.parent, s.stamp, s.inq, s.srcid, s.hdist, m.pos, dSlots, dTargets, 1, V);
        {
            DevGraph none;
            SolveArgs args = make_args(s, none, m);
            hipLaunchKernelGGL(k_admit, dim3(1), dim3(64), 0, 0, args, dSlots, dKinds, dTargets, dWant, dGens, dOff,
                               dSeeds, stageW, s.delta, s.deltaMsad, dXoff, dXoff, stageW, s.hdist);
        }
        spHost[0].kind = Q_MSAD; spHost[0].target = -1; spHost[0].done = 0; spHost[0].wantParent = wantParent;
        spHost[0].gen = 1;
        for (;;) {
            int running = 0;
            std::string e = solver_round(s, g->d, m, running);
            if (!e.empty()) return svg_set_error(SVG_E_HIP, e);
            if (running == 0) break;
        }
        std::string e = solver_fetch(s, 0, dist, nullptr, src_id, parent);
        if (!e.empty()) return svg_set_error(SVG_E_HIP, e);
        return SVG_OK;
    });
}

// Surface-point queries (SVG.cu:253-352) on the graph: the ICH runs and
// direct answers of up to `chunk` queries at once, then the graph stage of
// those without a source-side key point through the solver's slots (SVG.cu:
// 284-352), each joined on the device as its slot finishes.  For one query,
// path (cap entries) receives the graph vertices from last_vert back.
static int point_queries(svg_graph_s* g, int nq, const int32_t* faces, const double* pts, PointRes* out,
                         int32_t* path, int path_cap, int* path_len)
{
    const DevMesh& dm = g->mesh->d;
    const HostMesh& hm = g->mesh->h;
    for (int i = 0; i < 2 * nq; ++i)
        if (faces[i] < 0 || faces[i] >= hm.nf) return svg_set_error(SVG_E_ARG, "face out of range");
    for (int i = 0; i < 6 * nq; ++i)
        if (!std::isfinite(pts[i])) return svg_set_error(SVG_E_ARG, "point not finite");
    int rc = svg_mesh_upload_impl(const_cast<svg_mesh_s*>(g->mesh), true);
    if (rc) return rc;
    // a chunk's tables (about 570 KB per query at K = 500) within 8 GB; larger
    // chunks keep more propagations in flight (8,192 queries: 5.8K/s against
    // 4.6K/s in chunks of 2,048)
    const size_t perQuery = point_bytes_per_query(g->d.K);
    const int chunk = (int)std::max<size_t>(1, std::min<size_t>({(size_t)nq, 8192, ((size_t)8 << 30) / perQuery}));
    if (!g->point) g->point = new PointWork();
    std::string e = point_alloc(*g->point, g->d.K, chunk);
    if (!e.empty()) return svg_set_error(SVG_E_MEM, e);
    PointWork& w = *g->point;
    const int hopCap = 4 * (hm.nf + hm.nv) + 16;
    // graph stages share the solver's rounds: 256 slots (2 MB each at V = 72K)
    const int S = std::min(256, std::max(1, nq));
    SolverState* st = nullptr;
    rc = get_solver(g, S, &st);
    if (rc) return rc;
    if (path_len) *path_len = 0;
    for (int c0 = 0; c0 < nq; c0 += chunk) {
        const int n = std::min(chunk, nq - c0);
        e = point_ich(w, dm, n, faces + 2 * (size_t)c0, pts + 6 * (size_t)c0, hopCap, st->stream, out + c0);
        if (!e.empty()) return svg_set_error(SVG_E_HIP, e);
        // a run that stalled (a point on a vertex of its face) answers
        // nothing: route -2, distance NaN
        std::vector<int> todo;
        std::vector<int> stalled;
        for (int q = 0; q < n; ++q) {
            const uint32_t stt = out[c0 + q].status;
            if ((stt & SF_STALL_P0) || (out[c0 + q].srcEdge == -1 && (stt & SF_STALL_P1))) stalled.push_back(q);
            else if (out[c0 + q].srcEdge == -1) todo.push_back(q);
        }
        auto mark_stalled = [&]() {
            for (int q : stalled) {
                PointRes& r = out[c0 + q];
                r.dist = std::nan(""); r.route = -2;
                r.srcEdge = r.dstEdge = -1; r.srcX = r.dstX = 0.0; r.lastVert = r.firstVert = -1;
            }
        };
        if (todo.empty()) { mark_stalled(); continue; }
        std::vector<int> slotQuery((size_t)st->S, -1);
        size_t next = 0, done = 0;
        while (done < todo.size()) {
            std::vector<Admission> adm;
            for (int sl = 0; sl < st->S && next < todo.size(); ++sl) {
                if (slotQuery[sl] != -1) continue;
                const int q = todo[next++];
                Admission a;
                e = point_seeds(w, q, 0, st->stream, a.src, a.w);
                if (e.empty()) e = point_seeds(w, q, 1, st->stream, a.exitV, a.exitOff);
                if (!e.empty()) return svg_set_error(SVG_E_HIP, e);
                // the Dijkstra of SVG.cu:292, stopped once nothing left can
                // improve the join over the p1 run's vertices (same answer)
                a.slot = sl; a.kind = Q_JOIN; a.target = -1; a.wantParent = true;
                adm.push_back(std::move(a));
                slotQuery[sl] = q;
            }
            if (!adm.empty()) {
                e = solver_admit_many(*st, dm, adm);
                if (!e.empty()) return svg_set_error(SVG_E_ARG, e);
            }
            int running = 0;
            e = solver_round(*st, g->d, dm, running);
            if (!e.empty()) return svg_set_error(SVG_E_HIP, e);
            std::vector<int32_t> jobs;
            for (int sl = 0; sl < st->S; ++sl) {
                if (slotQuery[sl] == -1 || !st->spHost[sl].done) continue;
                jobs.push_back(slotQuery[sl]);
                jobs.push_back(sl);
            }
            e = point_join(w, dm, jobs, st->dist, st->parent, hopCap, st->stream);
            if (!e.empty()) return svg_set_error(SVG_E_HIP, e);
            for (int sl = 0; sl < st->S; ++sl) {
                if (slotQuery[sl] == -1 || !st->spHost[sl].done) continue;
                if (nq == 1 && path_cap > 0) {
                    PointRes r1;
                    e = point_results(w, 1, st->stream, &r1);
                    if (!e.empty()) return svg_set_error(SVG_E_HIP, e);
                    std::vector<int32_t> par((size_t)st->V);
                    e = solver_fetch(*st, sl, nullptr, nullptr, nullptr, par.data());
                    if (!e.empty()) return svg_set_error(SVG_E_HIP, e);
                    int k = 0;
                    for (int u = r1.lastVert; u >= 0 && u < st->V && k < path_cap; u = par[u]) path[k++] = u;
                    if (path_len) *path_len = k;
                }
                solver_release_slot(*st, sl);
                slotQuery[sl] = -1;
                ++done;
            }
        }
        e = point_results(w, n, st->stream, out + c0);
        if (!e.empty()) return svg_set_error(SVG_E_HIP, e);
        mark_stalled();
    }
    return SVG_OK;
}

extern "C" int svg_sssd_point(svg_graph_t g, int f0, const double* p0, int f1, const double* p1, double* dist,
                              int32_t* key_edge, double* key_x, int32_t* last_vert, int* route, int32_t* path,
                              int path_cap, int* path_len)
{
    return svg_guarded<int>([&]() -> int {
        if (!g || !p0 || !p1 || !dist) return svg_set_error(SVG_E_ARG, "null argument");
        if (path_cap > 0 && !path) return svg_set_error(SVG_E_ARG, "null path");
        std::lock_guard<std::mutex> lk(g->mu);
        const int32_t faces[2] = {f0, f1};
        const double pts[6] = {p0[0], p0[1], p0[2], p1[0], p1[1], p1[2]};
        PointRes r;
        int rc = point_queries(g, 1, faces, pts, &r, path, path_cap, path_len);
        if (rc) return rc;
        if (r.route == -2)
            return svg_set_error(SVG_E_ARG, "the propagation from a point stalled on unordered keys "
                                            "(a point on a vertex of its face)");
        *dist = r.dist;
        if (key_edge) { key_edge[0] = r.srcEdge; key_edge[1] = r.dstEdge; }
        if (key_x) { key_x[0] = r.srcX; key_x[1] = r.dstX; }
        if (last_vert) *last_vert = r.lastVert;
        if (route) *route = r.route;
        return SVG_OK;
    });
}

extern "C" int svg_sssd_point_batch(svg_graph_t g, int nq, const int32_t* faces, const double* points,
                                    double* dist, int32_t* key_edge, double* key_x, int32_t* last_vert,
                                    int32_t* route)
{
    return svg_guarded<int>([&]() -> int {
        if (!g || !faces || !points || !dist || nq < 0) return svg_set_error(SVG_E_ARG, "bad argument");
        if (nq == 0) return SVG_OK;
        std::lock_guard<std::mutex> lk(g->mu);
        std::vector<PointRes> r((size_t)nq);
        int rc = point_queries(g, nq, faces, points, r.data(), nullptr, 0, nullptr);
        if (rc) return rc;
        for (int q = 0; q < nq; ++q) {
            dist[q] = r[q].dist;
            if (key_edge) { key_edge[2 * q] = r[q].srcEdge; key_edge[2 * q + 1] = r[q].dstEdge; }
            if (key_x) { key_x[2 * q] = r[q].srcX; key_x[2 * q + 1] = r[q].dstX; }
            if (last_vert) last_vert[q] = r[q].lastVert;
            if (route) route[q] = r[q].route;
        }
        return SVG_OK;
    });
}

extern "C" int svg_msad_batch(svg_graph_t g, const int32_t* src, const double* w, const int64_t* offsets, int nq,
                              double* dist, int32_t* src_id)
{
    return svg_guarded<int>([&]() -> int {
        if (!g || !offsets || !dist || nq < 0) return svg_set_error(SVG_E_ARG, "bad argument");
        const size_t V = (size_t)g->d.V;
        for (int i = 0; i < nq; ++i) {
            if (offsets[i + 1] <= offsets[i]) return svg_set_error(SVG_E_ARG, "every query needs a source");
            if (offsets[i] < 0) return svg_set_error(SVG_E_ARG, "negative offset");
        }
        if (nq > 0 && (!src || !w)) return svg_set_error(SVG_E_ARG, "null sources or weights");
        for (int64_t i = nq > 0 ? offsets[0] : 0; nq > 0 && i < offsets[nq]; ++i)
            if (!std::isfinite(w[i]) || std::signbit(w[i])) return svg_set_error(SVG_E_ARG, kBadWeight);
        std::lock_guard<std::mutex> lk(g->mu);
        const int S = std::min(32, std::max(1, nq));
        SolverState* st = nullptr;
        int rc = get_solver(g, S, &st);
        if (rc) return rc;
        std::vector<int> slotQuery((size_t)st->S, -1);
        int nextQ = 0, done = 0;
        while (done < nq) {
            std::vector<Admission> adm;
            for (int q = 0; q < st->S && nextQ < nq; ++q) {
                if (slotQuery[q] != -1) continue;
                Admission a;
                a.slot = q; a.kind = Q_MSAD; a.target = -1; a.wantParent = src_id != nullptr;
                a.src.assign(src + offsets[nextQ], src + offsets[nextQ + 1]);
                a.w.assign(w + offsets[nextQ], w + offsets[nextQ + 1]);
                adm.push_back(std::move(a));
                slotQuery[q] = nextQ++;
            }
            std::string ea = solver_admit_many(*st, g->mesh->d, adm);
            if (!ea.empty()) return svg_set_error(SVG_E_ARG, ea);
            int running = 0;
            std::string e = solver_round(*st, g->d, g->mesh->d, running);
            if (!e.empty()) return svg_set_error(SVG_E_HIP, e);
            for (int q = 0; q < st->S; ++q) {
                if (slotQuery[q] == -1 || !st->spHost[q].done) continue;
                const size_t o = (size_t)slotQuery[q] * V;
                e = solver_fetch(*st, q, dist + o, nullptr, src_id ? src_id + o : nullptr, nullptr);
                if (!e.empty()) return svg_set_error(SVG_E_HIP, e);
                solver_release_slot(*st, q);
                slotQuery[q] = -1;
                ++done;
            }
        }
        return SVG_OK;
    });
}

extern "C" int svg_sssd_batch(svg_graph_t g, const int32_t* stp, int nq, double* out)
{
    return svg_guarded<int>([&]() -> int {
        if (!g || !stp || !out || nq < 0) return svg_set_error(SVG_E_ARG, "bad argument");
        std::lock_guard<std::mutex> lk(g->mu);
        const int S = std::min(64, std::max(1, nq));
        SolverState* st = nullptr;
        int rc = get_solver(g, S, &st);
        if (rc) return rc;
        std::vector<int> slotQuery((size_t)st->S, -1);
        int nextQ = 0, done = 0;
        const double w0 = 0.0;
        while (done < nq) {
            std::vector<Admission> adm;
            for (int q = 0; q < st->S && nextQ < nq; ++q) {
                if (slotQuery[q] != -1) continue;
                Admission a;
                a.slot = q; a.kind = Q_SSSD; a.target = stp[2 * nextQ + 1]; a.wantParent = false;
                a.src.push_back(stp[2 * nextQ]);
                a.w.push_back(w0);
                adm.push_back(std::move(a));
                slotQuery[q] = nextQ++;
            }
            std::string ea = solver_admit_many(*st, g->mesh->d, adm);
            if (!ea.empty()) return svg_set_error(SVG_E_ARG, ea);
            int running = 0;
            std::string e = solver_round(*st, g->d, g->mesh->d, running);
            if (!e.empty()) return svg_set_error(SVG_E_HIP, e);
            for (int q = 0; q < st->S; ++q) {
                if (slotQuery[q] == -1 || !st->spHost[q].done) continue;
                e = solver_fetch(*st, q, nullptr, &out[slotQuery[q]], nullptr, nullptr);
                if (!e.empty()) return svg_set_error(SVG_E_HIP, e);
                solver_release_slot(*st, q);
                slotQuery[q] = -1;
                ++done;
            }
        }
        return SVG_OK;
    });
}
