// Multi-query bucketed frontier engine (see solve.h) and the query entry
// points of the C ABI (svg_sssd, svg_msad, svg_sssd_batch).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cstring>
#include <vector>

#include "solve.h"
#include "svg_capi_internal.h"

namespace svg {

constexpr unsigned long long kInfBits = 0x7FEFFFFFFFFFFFFFull;   // DBL_MAX (GraphDistInfo default)
constexpr int kSolveBlock = 256;
constexpr int kSolveWaves = kSolveBlock / 64;

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
    const uint64_t* frontCur;
    unsigned long long* usedCur;   // per entry: the distance it was expanded with
    uint64_t* frontNext;
    int* expList;                  // entries to expand this round
    int* fcount;          // [cur, next] as selected by curIdx
    int curIdx;
    SlotParam* sp;
    uint32_t round;
    int frontCap;
    int* overflow;        // set when a frontier write would exceed frontCap
};

__device__ __forceinline__ double heur(const double* __restrict__ pos, int v, int t)
{
    // (verts[v].pos - verts[t].pos).length()  (SVG.cu:402)
    const double x = pos[4 * v] - pos[4 * t];
    const double y = pos[4 * v + 1] - pos[4 * t + 1];
    const double z = pos[4 * v + 2] - pos[4 * t + 2];
    return sqrt(x * x + y * y + z * z);
}

__global__ void k_slot_init(unsigned long long* dist, int32_t* parent, uint32_t* stamp, uint32_t* inq,
                            int32_t* srcid, const int* slots, int nslots, int V)
{
    const size_t total = (size_t)nslots * (size_t)V;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(i / V);
        const size_t o = (size_t)slots[k] * V + (i - (size_t)k * V);
        dist[o] = kInfBits;
        parent[o] = -1;
        stamp[o] = 0;
        inq[o] = 0;
        srcid[o] = -1;
    }
}

__device__ __forceinline__ uint64_t pack_entry(int gen, int slot, int v)
{
    return ((uint64_t)(uint16_t)gen << 48) | ((uint64_t)(uint16_t)slot << 32) | (uint64_t)(uint32_t)v;
}

// One round, pass 1: one lane per frontier entry.  Deferred entries move to
// the next frontier; entries to expand are listed (with the distance they
// expand with) for pass 2.
__global__ void __launch_bounds__(kSolveBlock) k_classify(SolveArgs a)
{
    const int lane = threadIdx.x & 63;
    const int n = min(a.fcount[a.curIdx], a.frontCap);
    int* nextCount = &a.fcount[a.curIdx ^ 1];
    int* expCount = &a.fcount[3];
    const int stride = gridDim.x * blockDim.x;
    for (int f0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63); f0 < n; f0 += stride) {
        const int f = f0 + lane;
        int q = 0, act = 0;           // act: 0 none, 1 defer, 2 expand
        bool hasDef = false;
        unsigned long long defKey = ~0ull;
        if (f < n) {
            const uint64_t e = a.frontCur[f];
            q = (int)((e >> 32) & 0xFFFFu);
            const int v = (int)(e & 0xFFFFFFFFu);
            const SlotParam& sp = a.sp[q];
            const size_t o = (size_t)q * a.V + v;
            if (sp.kind < 0 || (int)(e >> 48) != (sp.gen & 0xFFFF)) {
                // entry of an earlier query in this slot: dropped
            } else if (sp.done) {
                a.inq[o] = 0;
            } else {
                // Take the entry first, then read the distance: an improvement
                // that lands after the flag is cleared re-enters the frontier
                // itself, and one that landed before is visible in the read
                // (the acquire pairs with the improver's release exchange).
                __hip_atomic_exchange(&a.inq[o], 0u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
                const unsigned long long db =
                    __hip_atomic_load(&a.dist[o], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const double dv = __longlong_as_double((long long)db);
                double key = dv;
                bool prune = false;
                if (sp.kind == Q_SSSD) {
                    const double ub = __longlong_as_double((long long)__hip_atomic_load(
                        &a.dist[(size_t)q * a.V + sp.target], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
                    key = dv + heur(a.pos, v, sp.target);
                    prune = key * (1.0 - 1e-12) > ub;
                }
                if (prune) {
                    // no path through v can beat the answer any more
                } else if (key >= sp.thr) {
                    // deferred: re-arm the flag unless an improvement already
                    // put v back into the next frontier
                    if (__hip_atomic_exchange(&a.inq[o], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == 0u)
                        act = 1;
                    defKey = (unsigned long long)__double_as_longlong(key);
                    hasDef = true;
                } else {
                    act = 2;
                    a.usedCur[f] = db;
                }
            }
        }
        const unsigned long long dm = __ballot(act == 1);
        if (dm) {
            int basePos = 0;
            if (lane == 0) basePos = atomicAdd(nextCount, __popcll(dm));
            basePos = __shfl(basePos, 0);
            const int dpos = basePos + __popcll(dm & ((1ull << lane) - 1ull));
            if (act == 1) {
                if (dpos < a.frontCap) a.frontNext[dpos] = a.frontCur[f];
                else atomicOr(a.overflow, 1);
            }
        }
        const unsigned long long em = __ballot(act == 2);
        if (em) {
            int basePos = 0;
            if (lane == 0) basePos = atomicAdd(expCount, __popcll(em));
            basePos = __shfl(basePos, 0);
            if (act == 2) a.expList[basePos + __popcll(em & ((1ull << lane) - 1ull))] = f;
        }
        // One update per slot present in this wave instead of one per entry:
        // the per-slot counters are single addresses every lane would hit.
        unsigned long long todo = __ballot(hasDef || act == 2);
        while (todo) {
            const int leader = __ffsll((long long)todo) - 1;
            const int lq = __shfl(q, leader);
            const bool mine = (hasDef || act == 2) && q == lq;
            const unsigned long long grp = __ballot(mine);
            todo &= ~grp;
            unsigned long long k = mine && hasDef ? defKey : ~0ull;
            for (int off = 32; off > 0; off >>= 1) {
                const unsigned long long o = __shfl_xor(k, off);
                k = o < k ? o : k;
            }
            const bool anyExp = __ballot(mine && act == 2) != 0ull;
            if (lane == leader) {
                if (k != ~0ull) atomicMin(&a.sp[lq].minDeferred, k);
                if (anyExp) a.sp[lq].expanded = 1;
            }
        }
    }
}

// One round, pass 2: one wave per listed entry, 64 arcs of its row per step.
// A plain read of the target's distance filters the relaxations first; it
// can only be stale-high (distances only fall), so no improvement is missed.
__global__ void __launch_bounds__(kSolveBlock) k_expand(SolveArgs a)
{
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nw = (gridDim.x * blockDim.x) >> 6;
    const int n = a.fcount[3];
    int* nextCount = &a.fcount[a.curIdx ^ 1];
    for (int k = wave; k < n; k += nw) {
        const int f = a.expList[k];
        const uint64_t e = a.frontCur[f];
        const int qq = (int)((e >> 32) & 0xFFFFu);
        const int vv = (int)(e & 0xFFFFFFFFu);
        const double dvv = __longlong_as_double((long long)a.usedCur[f]);
        const SlotParam& sp = a.sp[qq];
        const int kind = sp.kind;
        const int t = sp.target;
        const int wantParent = sp.wantParent;
        const int gen = sp.gen;
        const size_t sb = (size_t)qq * a.V;
        const int tl = a.tail[vv];
        const int tlr = (tl + 63) & ~63;
        const size_t row = (size_t)vv * a.K;
        double ub = DBL_MAX;
        if (kind == Q_SSSD) ub = __longlong_as_double((long long)a.dist[sb + t]);
        for (int i = lane; i < tlr; i += 64) {
            bool push = false;
            int x = -1;
            if (i < tl) {
                x = a.adj[row + i];
                const double nd = dvv + a.w[row + i];
                const unsigned long long nb = (unsigned long long)__double_as_longlong(nd);
                bool ok = nb < a.dist[sb + x];
                if (ok && kind == Q_SSSD) ok = !((nd + heur(a.pos, x, t) * (1.0 - 1e-12)) > ub);
                if (ok) {
                    const unsigned long long old = atomicMin(&a.dist[sb + x], nb);
                    if (nb < old) {
                        a.stamp[sb + x] = a.round;
                        if (wantParent) a.parent[sb + x] = INT_MAX;
                        push = __hip_atomic_exchange(&a.inq[sb + x], 1u, __ATOMIC_ACQ_REL,
                                                     __HIP_MEMORY_SCOPE_AGENT) == 0u;
                    }
                }
            }
            const unsigned long long pm = __ballot(push);
            if (pm) {
                int basePos = 0;
                if (lane == 0) basePos = atomicAdd(nextCount, __popcll(pm));
                basePos = __shfl(basePos, 0);
                const int ppos = basePos + __popcll(pm & ((1ull << lane) - 1ull));
                if (push) {
                    if (ppos < a.frontCap) a.frontNext[ppos] = pack_entry(gen, qq, x);
                    else atomicOr(a.overflow, 1);
                }
            }
        }
        if (lane == 0) {
            atomicAdd((unsigned long long*)&a.sp[qq].relaxed, (unsigned long long)tl);
            atomicAdd((unsigned long long*)&a.sp[qq].settled, 1ull);
        }
    }
}

// Parents of the vertices improved this round: the smallest u whose
// expansion value reproduces the new distance exactly.
__global__ void __launch_bounds__(kSolveBlock) k_parent_fix(SolveArgs a)
{
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nw = (gridDim.x * blockDim.x) >> 6;
    const int n = a.fcount[3];
    for (int k = wave; k < n; k += nw) {
        const int f = a.expList[k];
        const uint64_t e = a.frontCur[f];
        const int qq = (int)((e >> 32) & 0xFFFFu);
        if (!a.sp[qq].wantParent) continue;
        const int vv = (int)(e & 0xFFFFFFFFu);
        const double dvv = __longlong_as_double((long long)a.usedCur[f]);
        const size_t sb = (size_t)qq * a.V;
        const int tl = a.tail[vv];
        const size_t row = (size_t)vv * a.K;
        for (int i = lane; i < tl; i += 64) {
            const int x = a.adj[row + i];
            if (a.stamp[sb + x] != a.round) continue;
            const double nd = dvv + a.w[row + i];
            if ((unsigned long long)__double_as_longlong(nd) == a.dist[sb + x])
                atomicMin(&a.parent[sb + x], vv);
        }
    }
}

__global__ void k_bookkeep(SolveArgs a, int S)
{
    for (int q = threadIdx.x; q < S; q += blockDim.x) {
        SlotParam& sp = a.sp[q];
        if (sp.kind < 0 || sp.done) continue;
        sp.rounds++;
        if (sp.expanded == 0) {
            if (sp.minDeferred == 0xFFFFFFFFFFFFFFFFull) {
                sp.done = 1;
            } else {
                const double md = __longlong_as_double((long long)sp.minDeferred);
                bool finished = false;
                if (sp.kind == Q_SSSD) {
                    const double ub = __longlong_as_double((long long)a.dist[(size_t)q * a.V + sp.target]);
                    finished = md * (1.0 - 1e-12) > ub;
                }
                if (finished) sp.done = 1;
                else sp.thr = md + sp.delta;
            }
        }
        sp.expanded = 0;
        sp.minDeferred = 0xFFFFFFFFFFFFFFFFull;
    }
    if (threadIdx.x == 0) { a.fcount[a.curIdx] = 0; a.fcount[3] = 0; }
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
    double* ds = nullptr;
    unsigned long long* dc = nullptr;
    hipMalloc(&ds, sizeof(double));
    hipMalloc(&dc, sizeof(unsigned long long));
    hipMemset(ds, 0, sizeof(double));
    hipMemset(dc, 0, sizeof(unsigned long long));
    hipLaunchKernelGGL(k_mean_arc, dim3(256), dim3(256), 0, 0, g.tail, g.w, g.V, g.K, ds, dc);
    double s = 0.0;
    unsigned long long c = 0;
    hipMemcpy(&s, ds, sizeof(double), hipMemcpyDeviceToHost);
    hipMemcpy(&c, dc, sizeof(unsigned long long), hipMemcpyDeviceToHost);
    hipFree(ds); hipFree(dc);
    return c ? s / (double)c : 1.0;
}

static SolveArgs make_args(SolverState& s, const DevGraph& g, const DevMesh& m)
{
    SolveArgs a;
    a.V = s.V; a.K = s.K;
    a.tail = g.tail; a.adj = g.adj; a.w = g.w; a.pos = m.pos;
    a.dist = s.dist; a.parent = s.parent; a.stamp = s.stamp; a.inq = s.inq; a.srcid = s.srcid;
    a.frontCur = s.front[s.cur]; a.usedCur = s.used[s.cur]; a.frontNext = s.front[s.cur ^ 1];
    a.fcount = s.fcount; a.curIdx = s.cur; a.sp = s.sp; a.round = s.round;
    a.frontCap = (int)s.frontCap; a.overflow = s.fcount + 2; a.expList = s.expList;
    return a;
}

#define HTRY(x)                                                                   \
    do {                                                                          \
        hipError_t _e = (x);                                                      \
        if (_e != hipSuccess) return std::string(#x ": ") + hipGetErrorString(_e); \
    } while (0)

std::string solver_create(const DevGraph& g, const DevMesh& m, int slots, double delta, SolverState& s)
{
    (void)m;
    if (slots < 1 || slots > 4096) return "slots must be in [1, 4096]";
    s.V = g.V; s.K = g.K; s.S = slots; s.delta = delta;
    const size_t sv = (size_t)slots * (size_t)g.V;
    s.frontCap = 2 * sv + 65536;
    if (s.frontCap >= 0x7fffffffull) return "slots * V too large for the frontier";
    HTRY(hipMalloc(&s.dist, sizeof(unsigned long long) * sv));
    HTRY(hipMalloc(&s.parent, sizeof(int32_t) * sv));
    HTRY(hipMalloc(&s.stamp, sizeof(uint32_t) * sv));
    HTRY(hipMalloc(&s.inq, sizeof(uint32_t) * sv));
    HTRY(hipMalloc(&s.srcid, sizeof(int32_t) * sv));
    for (int i = 0; i < 2; ++i) {
        HTRY(hipMalloc(&s.front[i], sizeof(uint64_t) * s.frontCap));
        HTRY(hipMalloc(&s.used[i], sizeof(unsigned long long) * s.frontCap));
    }
    HTRY(hipMalloc(&s.expList, sizeof(int) * s.frontCap));
    HTRY(hipMalloc(&s.fcount, sizeof(int) * 4));
    HTRY(hipMemset(s.fcount, 0, sizeof(int) * 4));
    HTRY(hipMalloc(&s.sp, sizeof(SlotParam) * (size_t)slots));
    HTRY(hipHostMalloc(&s.spHost, sizeof(SlotParam) * (size_t)slots, hipHostMallocDefault));
    std::vector<SlotParam> init((size_t)slots);
    for (auto& p : init) { std::memset(&p, 0, sizeof(p)); p.kind = -1; p.done = 1; p.minDeferred = ~0ull; }
    HTRY(hipMemcpy(s.sp, init.data(), sizeof(SlotParam) * (size_t)slots, hipMemcpyHostToDevice));
    std::memcpy(s.spHost, init.data(), sizeof(SlotParam) * (size_t)slots);
    s.cur = 0;
    s.round = 1;
    s.gens.assign((size_t)slots, 0);
    return std::string();
}

void solver_destroy(SolverState& s)
{
    hipFree(s.dist); hipFree(s.parent); hipFree(s.stamp); hipFree(s.inq); hipFree(s.srcid);
    for (int i = 0; i < 2; ++i) { hipFree(s.front[i]); hipFree(s.used[i]); }
    hipFree(s.fcount); hipFree(s.expList); hipFree(s.sp); hipHostFree(s.spHost); hipFree(s.stage); hipHostFree(s.stageHost);
    s = SolverState();
}

// Seeds of the admitted slots, one block each, in order (SVG.cu:246-249,
// 359-365): a repeated source keeps its last weight, as the reference's
// array writes do.  The first bucket bound is the smallest seed key + delta.
__global__ void k_admit(SolveArgs a, const int* __restrict__ slots, const int* __restrict__ kinds,
                        const int* __restrict__ targets, const int* __restrict__ wantP,
                        const int* __restrict__ gens, const int* __restrict__ off,
                        const int* __restrict__ seeds, const double* __restrict__ wts, double delta)
{
    if (threadIdx.x != 0) return;
    const int k = blockIdx.x;
    const int slot = slots[k], kind = kinds[k], target = targets[k], gen = gens[k];
    const size_t base = (size_t)slot * a.V;
    uint64_t* front = (uint64_t*)a.frontCur;
    double mk = DBL_MAX;
    for (int i = off[k]; i < off[k + 1]; ++i) {
        const int v = seeds[i];
        const double w = wts[i];
        a.dist[base + v] = (unsigned long long)__double_as_longlong(w);
        a.parent[base + v] = -1;
        a.srcid[base + v] = v;
        if (a.inq[base + v] == 0) {
            a.inq[base + v] = 1;
            const int pos = atomicAdd(&a.fcount[a.curIdx], 1);
            if (pos < a.frontCap) front[pos] = pack_entry(gen, slot, v);
            else atomicOr(a.overflow, 1);
        }
        const double key = kind == Q_SSSD ? w + heur(a.pos, v, target) : w;
        mk = key < mk ? key : mk;
    }
    SlotParam p;
    p.kind = kind; p.target = target; p.wantParent = wantP[k]; p.done = 0;
    p.thr = mk + delta; p.delta = delta; p.minDeferred = ~0ull;
    p.expanded = 0; p.rounds = 0; p.gen = gen; p.pad = 0; p.relaxed = 0; p.settled = 0;
    a.sp[slot] = p;
}

std::string solver_admit_many(SolverState& s, const DevMesh& m, const std::vector<Admission>& adm)
{
    const int n = (int)adm.size();
    if (n == 0) return std::string();
    HTRY(hipStreamSynchronize(0));      // the pinned staging buffer is free again
    std::vector<int> ints;          // slots, kinds, targets, wantP, gens, off[n+1], seeds
    ints.reserve((size_t)6 * n + 64);
    size_t nseeds = 0;
    for (const Admission& a : adm) {
        if (a.slot < 0 || a.slot >= s.S) return "slot out of range";
        if (a.src.empty() || a.src.size() != a.w.size()) return "bad seeds";
        for (int v : a.src) if (v < 0 || v >= s.V) return "source out of range";
        if (a.kind == Q_SSSD && (a.target < 0 || a.target >= s.V)) return "target out of range";
        nseeds += a.src.size();
    }
    for (const Admission& a : adm) ints.push_back(a.slot);
    for (const Admission& a : adm) ints.push_back(a.kind);
    for (const Admission& a : adm) ints.push_back(a.kind == Q_SSSD ? a.target : -1);
    for (const Admission& a : adm) ints.push_back(a.wantParent ? 1 : 0);
    for (const Admission& a : adm) {
        s.gens[a.slot] = (s.gens[a.slot] + 1) & 0xFFFF;
        ints.push_back(s.gens[a.slot]);
    }
    int o = 0;
    for (const Admission& a : adm) { ints.push_back(o); o += (int)a.src.size(); }
    ints.push_back(o);
    for (const Admission& a : adm) ints.insert(ints.end(), a.src.begin(), a.src.end());
    const size_t intBytes = ((ints.size() * sizeof(int) + 15) / 16) * 16;
    const size_t bytes = intBytes + nseeds * sizeof(double);
    if (bytes > s.stageCap) {
        hipFree(s.stage);
        s.stageCap = std::max(bytes, (size_t)1 << 20);
        HTRY(hipMalloc(&s.stage, s.stageCap));
        hipHostFree(s.stageHost);
        HTRY(hipHostMalloc(&s.stageHost, s.stageCap, hipHostMallocDefault));
    }
    std::memcpy(s.stageHost, ints.data(), ints.size() * sizeof(int));
    double* wh = (double*)((char*)s.stageHost + intBytes);
    for (const Admission& a : adm) for (double w : a.w) *wh++ = w;
    HTRY(hipMemcpyAsync(s.stage, s.stageHost, bytes, hipMemcpyHostToDevice, 0));
    const int* d = (const int*)s.stage;
    const int* dSlots = d;
    const int* dKinds = d + n;
    const int* dTargets = d + 2 * n;
    const int* dWant = d + 3 * n;
    const int* dGens = d + 4 * n;
    const int* dOff = d + 5 * n;
    const int* dSeeds = d + 6 * n + 1;
    const double* dW = (const double*)((const char*)s.stage + intBytes);
    const size_t total = (size_t)n * s.V;
    hipLaunchKernelGGL(k_slot_init, dim3((unsigned)std::min<size_t>(2048, (total + 255) / 256)), dim3(256), 0, 0,
                       s.dist, s.parent, s.stamp, s.inq, s.srcid, dSlots, n, s.V);
    DevGraph dummy;
    SolveArgs args = make_args(s, dummy, m);
    hipLaunchKernelGGL(k_admit, dim3(n), dim3(64), 0, 0, args, dSlots, dKinds, dTargets, dWant, dGens, dOff,
                       dSeeds, dW, s.delta);
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

std::string solver_round(SolverState& s, const DevGraph& g, const DevMesh& m, int& running)
{
    SolveArgs a = make_args(s, g, m);
    s.anyParent = false;
    for (int q = 0; q < s.S; ++q)
        if (s.spHost[q].kind >= 0 && !s.spHost[q].done && s.spHost[q].wantParent) s.anyParent = true;
    hipLaunchKernelGGL(k_classify, dim3(512), dim3(kSolveBlock), 0, 0, a);
    hipLaunchKernelGGL(k_expand, dim3(2048), dim3(kSolveBlock), 0, 0, a);
    if (s.anyParent) hipLaunchKernelGGL(k_parent_fix, dim3(2048), dim3(kSolveBlock), 0, 0, a);
    hipLaunchKernelGGL(k_bookkeep, dim3(1), dim3(256), 0, 0, a, s.S);
    s.launches += s.anyParent ? 4 : 3;
    hipError_t e = hipMemcpy(s.spHost, s.sp, sizeof(SlotParam) * (size_t)s.S, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return std::string("solver round: ") + hipGetErrorString(e);
    int ov = 0;
    hipMemcpy(&ov, s.fcount + 2, sizeof(int), hipMemcpyDeviceToHost);
    if (ov) return "solver frontier overflow";
    s.cur ^= 1;
    s.round++;
    running = 0;
    for (int q = 0; q < s.S; ++q)
        if (s.spHost[q].kind >= 0 && !s.spHost[q].done) ++running;
    return std::string();
}

std::string solver_fetch(SolverState& s, int slot, double* distOut, double* distT, int32_t* srcidOut,
                         int32_t* parentOut)
{
    const size_t sb = (size_t)slot * s.V;
    if (distT) {
        unsigned long long b;
        HTRY(hipMemcpy(&b, s.dist + sb + s.spHost[slot].target, sizeof(b), hipMemcpyDeviceToHost));
        std::memcpy(distT, &b, sizeof(double));
    }
    if (distOut) HTRY(hipMemcpy(distOut, s.dist + sb, sizeof(double) * (size_t)s.V, hipMemcpyDeviceToHost));
    if (srcidOut) {
        hipLaunchKernelGGL(k_resolve_src, dim3((s.V + 255) / 256), dim3(256), 0, 0, s.parent, s.srcid, s.dist,
                           s.V, slot);
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
}

extern "C" int svg_msad(svg_graph_t g, const int32_t* src, const double* w, int ns, double* dist, int32_t* src_id,
                        int32_t* parent)
{
    if (!g || !src || !w || !dist || ns <= 0) return svg_set_error(SVG_E_ARG, "bad argument");
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
}

extern "C" int svg_sssd_batch(svg_graph_t g, const int32_t* stp, int nq, double* out)
{
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
}
