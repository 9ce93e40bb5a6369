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

#include <cstdint>
#include <string>
#include <vector>

#include "svg_graph.h"
#include "svg_mesh.h"

namespace svg {

enum : int { Q_SSSD = 0, Q_MSAD = 1 };

struct SlotParam {              // device-resident, one per slot
    int kind;                   // Q_SSSD / Q_MSAD, -1 = idle
    int target;                 // SSSD target
    int wantParent;             // keep parents (MSAD src ids, SSSD paths)
    int done;
    double thr;                 // current bucket bound (key < thr is expanded)
    double delta;
    unsigned long long minDeferred;   // bits of the smallest deferred key
    int expanded;               // entries expanded this round
    int rounds;
    int gen;                    // generation tag of this slot's frontier entries
    int pad;
    long long relaxed;          // arcs scanned
    long long settled;          // entries expanded
};

struct SolverState {
    int V = 0, K = 0, S = 0;
    double delta = 0.0;
    unsigned long long* dist = nullptr;   // S*V bits of double
    int32_t* parent = nullptr;            // S*V
    uint32_t* stamp = nullptr;            // S*V round of last improvement
    uint32_t* inq = nullptr;              // S*V frontier membership
    int32_t* srcid = nullptr;             // S*V
    uint64_t* front[2] = {nullptr, nullptr};   // entries: gen << 48 | slot << 32 | vertex
    unsigned long long* used[2] = {nullptr, nullptr};   // per entry: dist used when expanded (0 = deferred)
    int* fcount = nullptr;                // [0,1] frontier sizes, [2] overflow, [3] expand count
    int* expList = nullptr;               // entries expanded this round
    bool anyParent = false;               // some active slot keeps parents
    SlotParam* sp = nullptr;              // S (device)
    SlotParam* spHost = nullptr;          // S (pinned)
    void* stage = nullptr;                // admission staging (device)
    void* stageHost = nullptr;            // admission staging (pinned)
    size_t stageCap = 0;
    int cur = 0;
    uint32_t round = 1;
    size_t frontCap = 0;
    long long launches = 0;
    std::vector<int> gens;                // host copy of each slot's generation
};

std::string solver_create(const DevGraph& g, const DevMesh& m, int slots, double delta, SolverState& s);
void solver_destroy(SolverState& s);
struct Admission {
    int slot = 0, kind = Q_SSSD, target = -1;
    bool wantParent = false;
    std::vector<int> src;
    std::vector<double> w;
};
// Seeds several slots with one init launch and one seed launch (stream 0).
std::string solver_admit_many(SolverState& s, const DevMesh& m, const std::vector<Admission>& adm);
// Seeds a slot (device work queued on stream 0).
std::string solver_admit(SolverState& s, const DevMesh& m, int slot, int kind, const int* src,
                         const double* w, int ns, int target, bool wantParent);
// One round for every active slot; returns the number of slots still running.
std::string solver_round(SolverState& s, const DevGraph& g, const DevMesh& m, int& running);
// Results of a finished slot.
std::string solver_fetch(SolverState& s, int slot, double* distOut /*V or null*/, double* distT,
                         int32_t* srcidOut, int32_t* parentOut);
void solver_release_slot(SolverState& s, int slot);
double graph_mean_arc(const DevGraph& g);

}  // namespace svg
