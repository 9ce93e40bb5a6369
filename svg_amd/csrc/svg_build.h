// Host-side build entry points (used by the C ABI in capi.cpp).
#pragma once

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "svg_graph.h"
#include "svg_mesh.h"

namespace svg {

// What the batched autotuner has learned about one (mesh, K, coefficients):
// every batched build records each source's event count and its tables'
// peak occupancy, and the propagation time of the block shape it used.
struct TuneRecord {
    int K = 0;
    unsigned splitCoef = 0, vertCoef = 0;
    std::vector<int32_t> cost;     // per source: window pops + 3 x pseudo pops (time-weighted)
    int maxVert = 0, maxSplit = 0; // largest table occupancy seen
    int maxWinQ = 1023, maxPseudoQ = 1023;  // longest window / pseudo queue seen
    double blockMs[8] = {0.0};     // propagation time of each launch shape tried
};

struct BuildConfig {
    int K = 0;
    unsigned splitCoef = 10;     // SVG::SetParameters (kernel.cu:422)
    unsigned vertCoef = 2;
    int mode = 1;                // 0 = plain, 1 = batched
    // plain kernel launch shape (reference: 64 x 128, SVG.cuh:10-11)
    int plainBlocks = 64;
    int plainThreads = 128;
    // batched kernel knobs (0 = choose automatically)
    int lanes = 0;               // propagation lanes in flight
    int propBlock = 0;           // threads per block of the propagation kernel
    int chunk = 0;               // sources whose tables are resident at once
    int vertCap = 0;             // per-source vertex table slots (power of two)
    int splitCap = 0;            // per-source split table slots (power of two)
    int lanesPerWave = 0;        // source lanes per 64-wide wave of the propagation kernel
    int queueCap = 0;            // window / pseudo queue storage per lane (0: the reference's 1023)
    int sourceOrder = 0;         // 0 auto, 1 vertex id order, 2 predicted from the mesh's saddles
    int autotune = 0;
    TuneRecord* tune = nullptr;  // autotune: what earlier builds of this mesh recorded
    // uploads the mesh's reference-layout arrays (the plain path's) when a
    // batched build first needs them; empty string on success
    std::function<std::string()> ensurePlainMesh;
};

struct BuildStats {
    double seconds = 0.0;        // device time of the build kernels (HIP events)
    double propSeconds = 0.0;    // propagation kernels
    double extractSeconds = 0.0; // extraction kernels
    double redoSeconds = 0.0;    // plain-path recomputation of flagged sources
    double allocSeconds = 0.0;
    long long arcs = 0;
    int redone = 0;              // sources recomputed on the plain path
    int flaggedDrop = 0, flaggedOverflow = 0, flaggedError = 0;
    int lanes = 0, propBlock = 0, chunk = 0, vertCap = 0, splitCap = 0;
    int tuned = 0;               // 1: order, capacities and block came from the autotuner
    int lanesPerWave = 0;
    int queueCap = 0, pseudoCap = 0;
    int fallbackPlain = 0;       // 1: the mesh exceeds the batched form's limits; every source ran plain
    int sourceOrder = 0;         // order the sources ran in: 1 id, 2 predicted, 3 recorded event counts
    std::vector<long long> counters;   // summed event counters (CNT_NUM)
    std::vector<uint32_t> status;      // per-vertex status bits (SF_*)
};

std::string upload_mesh(const HostMesh& h, DevMesh& d, bool aos);
void free_dev_mesh(DevMesh& d);
std::string alloc_graph(int V, int K, DevGraph& g);
void free_graph(DevGraph& g);

std::string build_plain(const HostMesh& h, const DevMesh& d, const BuildConfig& c, DevGraph& g,
                        BuildStats& st, const int32_t* srcList, int nList);
std::string build_batched(const HostMesh& h, const DevMesh& d, const BuildConfig& c, DevGraph& g,
                          BuildStats& st);

}  // namespace svg
