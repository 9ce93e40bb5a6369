// Geodesic distance between two surface points (SVG::SolveSSSD(f0, p0, f1,
// p1), SVG.cu:253-352): the device work buffers and the two launches around
// the graph stage.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "svg_mesh.h"

namespace svg {

constexpr int kStoredWin = 100;      // STORED_WIN_BUF_SIZE (SVG.cuh:15)

// The query's result (the SVGNode fields SolveSSSD writes, SVG.cu:253-352).
struct PointRes {
    double dist;             // geodDist
    double srcX, dstX;       // nextToSrcX, nextToDstX
    int srcEdge, dstEdge;    // nextToSrcEdge, nextToDstEdge
    int lastVert;            // saddleVertNearDst: last graph vertex (-1: direct)
    int firstVert;           // first graph vertex (-1: direct)
    int route;               // 0 direct, 1 through the graph, -1 nothing joins the points
    uint32_t status;         // SF_* bits of the two ICH runs
};

struct PointWork {
    int K = 0, cap = 0;                  // K of the graph; queries the buffers hold
    unsigned ssz = 0, vsz = 0;
    void* sp = nullptr;                  // split tables (K*10 + 1 records), 2 runs per query
    void* vt = nullptr;                  // vertex tables (K*2 + 1 records)
    void* wq = nullptr;                  // window heaps
    void* pq = nullptr;                  // pseudo-source heaps
    void* stored = nullptr;              // windows kept for the destination face, per query
    PointRes* res = nullptr;             // device results, per query
    int32_t* fDev = nullptr;             // faces (f0, f1) per query
    double* pDev = nullptr;              // points (p0, p1) per query
    int32_t* jobsDev = nullptr;          // (query, slot) pairs of a join launch
    int32_t* jobsHost = nullptr;         // pinned staging of the pairs
    int ldsOk = -1;                      // a run's heaps and vertex table fit a block's LDS (-1: not checked)
    std::vector<uint8_t> hostVt;         // host copy of one p0 vertex table
};

// Device bytes one query of the batch holds.
size_t point_bytes_per_query(int K);
std::string point_alloc(PointWork& w, int K, int cap);
void point_free(PointWork& w);

// Both ICH runs of n queries (faces[2n] = f0, f1; pts[6n] = p0, p1) and each
// query's direct answer through f1, copied to out[n].
std::string point_ich(PointWork& w, const DevMesh& m, int n, const int32_t* faces, const double* pts, int hopCap,
                      hipStream_t stream, PointRes* out);

// Every vertex the run of query q from side 0 (p0: the graph seeds,
// SVG.cu:283-291) or side 1 (p1: the exits the join scans) reached, with its
// distance, in its table's slot order.
std::string point_seeds(PointWork& w, int q, int side, hipStream_t stream, std::vector<int>& src,
                        std::vector<double>& dist);

// Joins finished graph stages with their p1 runs and backtraces both ends,
// into the device results (asynchronous on stream).  jobs = (query, solver
// slot) pairs; gdist / gparent are the solver's arrays (slot-major, V each).
std::string point_join(PointWork& w, const DevMesh& m, const std::vector<int32_t>& jobs,
                       const unsigned long long* gdist, const int32_t* gparent, int hopCap, hipStream_t stream);

// Copies the first n device results to out.
std::string point_results(PointWork& w, int n, hipStream_t stream, PointRes* out);

}  // namespace svg
