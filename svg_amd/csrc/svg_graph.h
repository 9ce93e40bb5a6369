// Device SVG in padded-row (ELL) form: row u holds tail[u] arcs at
// [u*K, u*K + tail[u]) of each SoA array.  This is the reference's K-stride
// row layout (SVG.cu:42, 136) split into separate arrays, so queries stream
// only adj and w.
#pragma once

#include <cstdint>

namespace svg {

struct DevGraph {
    int V = 0, K = 0;
    int32_t* tail = nullptr;      // V
    int32_t* adj = nullptr;       // V*K
    double* w = nullptr;          // V*K  geodesic length (SVGNode::geodDist)
    int32_t* srcEdge = nullptr;   // V*K  SVGNode::nextToSrcEdge
    int32_t* dstEdge = nullptr;   // V*K  SVGNode::nextToDstEdge
    double* srcX = nullptr;       // V*K  SVGNode::nextToSrcX
    double* dstX = nullptr;       // V*K  SVGNode::nextToDstX
};

// Per-source status bits written by the build kernels.
enum : uint32_t {
    SF_WIN_DROP = 1,        // a window push met a full heap (dropped, as the reference does)
    SF_PSEUDO_DROP = 2,
    SF_REF_OVERFLOW = 4,    // more distinct keys than the reference tables hold
    SF_CAPACITY = 8,        // the batched tables ran out of room: recomputed on the plain path
    SF_ERROR = 16,          // hop cap / unreachable branch
    SF_REDONE = 32,         // row produced by the exact plain path after a batched flag
    SF_STALL_P0 = 64,       // point query: the p0 run stopped on unordered (NaN) keys
    SF_STALL_P1 = 128,      // point query: the p1 run did
};

}  // namespace svg
