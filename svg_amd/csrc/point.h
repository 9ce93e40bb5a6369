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
    int K = 0;
    unsigned ssz = 0, vsz = 0;
    void* sp[2] = {nullptr, nullptr};    // split tables (K*10 + 1 records)
    void* vt[2] = {nullptr, nullptr};    // vertex tables (K*2 + 1 records)
    void* wq[2] = {nullptr, nullptr};    // window heaps
    void* pq[2] = {nullptr, nullptr};    // pseudo-source heaps
    void* stored = nullptr;              // windows kept for the destination face
    PointRes* res = nullptr;             // device result
    std::vector<uint8_t> hostVt;         // host copy of the p0 vertex table
};

std::string point_alloc(PointWork& w, int K);
void point_free(PointWork& w);

// Both ICH runs (p0 keeping the windows into f1, and p1) and the direct
// answer through f1; the result lands in `out`.  Points are xyz triples.
std::string point_ich(PointWork& w, const DevMesh& m, int f0, const double* p0, int f1, const double* p1,
                      int hopCap, hipStream_t stream, PointRes& out);

// Graph seeds from the p0 run: every vertex it reached, in its table's slot
// order (SVG.cu:283-291).
std::string point_seeds(PointWork& w, hipStream_t stream, std::vector<int>& src, std::vector<double>& dist);

// Joins the graph stage (dist bits / parents of V vertices, device) with the
// p1 run, then backtraces both ends.  `out` carries the direct attempt's
// result in and the final one out.
std::string point_join(PointWork& w, const DevMesh& m, const unsigned long long* gdist, const int32_t* gparent,
                       int f0, int f1, int hopCap, hipStream_t stream, PointRes& out);

}  // namespace svg
