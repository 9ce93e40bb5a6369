// Multi-query bucketed frontier engine on a device-resident SVG.
//
// Replaces SVG::SolveSSSD / SolveMSAD / Astar (SVG.cu:241-251, 354-413).
// Up to S queries ("slots") advance together: one launch expands the
// frontier entries of every active slot.  Distances are the least fixed point
// of d(v) = min(d(v), d(u) + w(u,v)); IEEE addition of non-negative weights is
// monotone, so any relaxation order reaches the same bits as the reference's
// heap Dijkstra.  A* queries prune a relaxation only when even the Euclidean
// lower bound to the target (SVG.cu:402) cannot beat the current answer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "svg_graph.h"
#include "svg_mesh.h"

namespace svg {

// Q_JOIN: a multi-source Dijkstra that only has to be exact where it meets a
// set of exit vertices with offsets (surface-point queries): it stops once no
// frontier key can beat the best exit distance + offset found so far.
enum : int { Q_SSSD = 0, Q_MSAD = 1, Q_JOIN = 2 };

struct SlotParam {              // device-resident, one per slot
    int kind;                   // Q_SSSD / Q_MSAD, -1 = idle
    int target;                 // SSSD target
    int wantParent;             // keep parents (MSAD src ids, SSSD paths)
    int done;
    double thr;                 // current bucket bound (key < thr is expanded)
    double delta;
    unsigned long long minDeferred;   // bits of the smallest deferred key (last round)
    int expanded;               // anything expanded in the last round
    int rounds;
    int gen;                    // generation tag of this slot's frontier entries
    int pad;
    long long relaxed;          // arcs scanned
    long long settled;          // entries expanded
    double result;              // SSSD: distance to the target at the last bookkeeping;
                                // JOIN: best exit distance + offset so far (an upper bound)
};

// Distances and keys are compared and lowered as the bits of non-negative
// doubles (atomicMin on their 64-bit patterns), so MSAD source weights must
// be finite with the sign bit clear; the reference takes any weight
// (SVG.cu:361), and the callers refuse the others with this message.
constexpr const char* kBadWeight = "source weight must be finite and non-negative (not -0.0)";
constexpr int kCBlocks = 512;     // classify blocks = deferred-list regions
constexpr int kPLists = 8;        // push lists (and won lists)
constexpr int kCntPad = 32;       // counters one 128-B line apart
constexpr int kAccCopies = 16;    // copies of the per-slot statistics

// Per-slot statistics of a round, spread over kAccCopies copies so the
// atomics of one slot do not all meet at one address.
struct SlotAcc {
    unsigned long long minDeferred;   // smallest key deferred by classify
    unsigned long long minImp;        // smallest key a relaxation lowered a target to
    unsigned long long settled;
    unsigned long long relaxed;
    int expanded;
    int pad;
};

// A row to relax: the frontier entry and the distance it expands with.
struct ExpEnt {
    uint64_t entry;
    unsigned long long dist;
};

// An improvement (or exact tie) found by a relaxation, kept so the round's
// parents can be resolved without rescanning the rows.
struct WonEnt {
    int32_t slot, x, u, pad;
    unsigned long long nb;
};

// Device-resident list sizes.  Index [p] is a round parity.
struct SolveCounts {
    int dcnt[2][kCBlocks];            // deferred entries per classify region
    int dstride[2];                   // region stride of the deferred lists
    int pcnt[2][kPLists * kCntPad];   // push-list sizes (padded)
    int ecnt[2][kCBlocks];            // rows per classify region
    int estride[2];
    int wcnt[2][kPLists * kCntPad];   // won-list sizes (padded)
    int overflow;
};

// Slots up to this count are bookkept inside the next round's classify pass
// (one LDS table per block); more slots use a bookkeeping pass per round.
constexpr int kLdsSlots = 256;

struct SolverState {
    int V = 0, K = 0, S = 0;
    double delta = 0.0;                   // bucket width of SSSD queries
    double deltaMsad = 0.0;               // bucket width of MSAD queries
    int roundsPerSync = 4;                // rounds queued between two host checks
    hipStream_t stream = 0;               // every launch and copy of this state (set before create)
    int chainBudget = 4;                  // same-bucket rows a wave expands itself per expand pass
    int expandBlocks = 768;               // expand grid: 768 = 3 waves per SIMD, the pass's occupancy;
                                          // states sharing the device concurrently take fewer
    unsigned long long* dist = nullptr;   // S*V bits of double
    int32_t* parent = nullptr;            // S*V
    uint32_t* stamp = nullptr;            // S*V round of last improvement
    uint32_t* inq = nullptr;              // S*V frontier membership
    int32_t* srcid = nullptr;             // S*V
    float* hdist = nullptr;               // S*V straight-line distance to the SSSD target (rounded down);
                                          // JOIN: exit offsets rounded up, +inf off the exit set
    uint64_t* dl[2] = {nullptr, nullptr};   // deferred lists (kCBlocks regions), by parity
    uint64_t* pl[2] = {nullptr, nullptr};   // kPLists push lists, by parity
    ExpEnt* exp[2] = {nullptr, nullptr};  // rows of a round (kCBlocks regions), by parity
    SolveCounts* cnt = nullptr;
    int* ovDev = nullptr;                 // overflow flag: the record after the last slot's
    size_t frontCap = 0;                  // capacity of a deferred / row list
    size_t pCap = 0;                      // capacity of one push list
    WonEnt* won[2] = {nullptr, nullptr};  // kPLists won lists of wonCap entries, by parity
    size_t wonCap = 0;
    SlotAcc* acc[2] = {nullptr, nullptr}; // kAccCopies * S statistics of a round, by parity
    double* thr = nullptr;                // 2 * S bucket bound of a round, by parity
    int* doneP = nullptr;                 // 2 * S finished flag of a round, by parity
    bool anyParent = false;               // some active slot keeps parents
    SlotParam* sp = nullptr;              // S + 1 (device; the last record carries ovDev)
    SlotParam* spHost = nullptr;          // S + 1 (pinned)
    void* stage[2] = {nullptr, nullptr};        // admission staging (device), alternating
    void* stageHost[2] = {nullptr, nullptr};    // admission staging (pinned)
    size_t stageCap[2] = {0, 0};
    hipEvent_t stageEv[2] = {nullptr, nullptr}; // staging copy done
    int stageIdx = 0;
    int cur = 0;                          // parity of the next round's input
    uint32_t round = 1;
    long long launches = 0;
    std::vector<int> gens;                // host copy of each slot's generation
};

// Device bytes one slot costs (for admission by memory).
size_t solver_bytes_per_slot(int V);
std::string solver_create(const DevGraph& g, const DevMesh& m, int slots, double delta, SolverState& s);
void solver_destroy(SolverState& s);
struct Admission {
    int slot = 0, kind = Q_SSSD, target = -1;
    bool wantParent = false;
    std::vector<int> src;
    std::vector<double> w;
    std::vector<int> exitV;         // Q_JOIN: exit vertices and their offsets
    std::vector<double> exitOff;
};
// Seeds several slots with one init launch and one seed launch (on s.stream).
std::string solver_admit_many(SolverState& s, const DevMesh& m, const std::vector<Admission>& adm);
// Seeds a slot (device work queued on s.stream).
std::string solver_admit(SolverState& s, const DevMesh& m, int slot, int kind, const int* src,
                         const double* w, int ns, int target, bool wantParent);
// Queues nRounds rounds for every active slot on s.stream, ending with the
// bookkeeping pass, so the slot records are current when they complete.  A
// slot that finishes early idles through the remaining rounds.
std::string solver_enqueue_rounds(SolverState& s, const DevGraph& g, const DevMesh& m, int nRounds);
// Queues a copy of the S + 1 slot records (the last one's `kind` is the
// overflow flag) into pinned memory.
std::string solver_enqueue_status(SolverState& s, SlotParam* hostDst);
// Queues, on `stream`, the results of a finished slot into pinned buffers:
// distances, nearest-source ids (resolved from the parent chains first),
// parents.  The caller orders `stream` after the rounds that finished it.
std::string solver_enqueue_fetch(SolverState& s, int slot, hipStream_t stream, double* distDst,
                                 int32_t* srcidDst, int32_t* parentDst);
// nRounds rounds for every active slot, queued back to back, then one host
// check; returns the number of slots still running.
std::string solver_rounds(SolverState& s, const DevGraph& g, const DevMesh& m, int nRounds, int& running);
inline std::string solver_round(SolverState& s, const DevGraph& g, const DevMesh& m, int& running)
{
    return solver_rounds(s, g, m, s.roundsPerSync, running);
}
// Results of a finished slot.
std::string solver_fetch(SolverState& s, int slot, double* distOut /*V or null*/, double* distT,
                         int32_t* srcidOut, int32_t* parentOut);
void solver_release_slot(SolverState& s, int slot);
double graph_mean_arc(const DevGraph& g);

}  // namespace svg
