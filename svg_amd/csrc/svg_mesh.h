// Host half-edge mesh and its device images.
//
// The half-edge rules are those of the reference loader (Mesh.cu:67-123):
// half-edge 3f+j runs tri[f][j] -> tri[f][(j+1)%3], next/prev stay inside the
// face, twins pair opposite half-edges, the vertex angle sum adds the corner
// angle at verts[0] of every half-edge in index order, and firstEdge is the
// largest-index outgoing half-edge.  On open meshes (which the reference
// cannot build) a boundary half-edge has twin -1, a boundary vertex's
// firstEdge is its outgoing half-edge whose prev has no twin, and boundary
// vertices are never saddles.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace svg {

// Vertex flag bits (device and host).
enum : uint8_t {
    VF_NOT_SADDLE = 1,   // angle < 2*PI, or boundary   (ICHDevice.cu:558, 766)
    VF_SADDLE = 2,       // angle > 2*PI, not boundary  (ICHDevice.cu:688)
    VF_BOUNDARY = 4,
};

// One triangle as the batched kernels read it: the three half-edges 3f+j
// with their lengths, origin vertices and twins.  64 bytes, one aligned read.
struct alignas(64) FaceRec {
    double l[3];
    int32_t v[3];
    int32_t tw[3];
    int32_t pad[4];
};
static_assert(sizeof(FaceRec) == 64, "FaceRec must be 64 bytes");

// The fan of a vertex v: its outgoing half-edges o_0 = firstEdge[v],
// o_{k+1} = next(twin(o_k)) (the order in which a pseudo-source visits its
// neighbours, ICHDevice.cu:775-791), one 64-byte record per position.  A
// pseudo-source's sweeps and windows (ICHDevice.cu:728-904) read these
// records by position instead of walking the fan one dependent read at a
// time: position k - 1 is twin(prev(o_k)) and the incoming half-edge
// twin(o_k) is followed by twin(o_{k+1}).
struct alignas(64) FanRec {
    double cA;        // cornerA(o): corner at v in face(o)
    double cB;        // cornerB(twin(o)): corner at v in face(twin(o)); 0 without a twin
    double lo;        // len(o)
    double lnx;       // len(next(o)): the edge opposite v in face(o)
    double lpv;       // len(prev(o))
    int32_t o;        // the half-edge
    int32_t twn;      // twin(next(o))
    int32_t nb;       // verts[1](o): the neighbour
    int32_t tw;       // twin(o)
    int32_t nbFlag;   // vertex flags of the neighbour
    int32_t pad;
};
static_assert(sizeof(FanRec) == 64, "FanRec must be 64 bytes");

// Reference-layout records for the plain kernel (Mesh.cuh:6-23).
struct EdgeAoS {
    int32_t v0, v1, twin, next, prev, face;
    double len;
};
static_assert(sizeof(EdgeAoS) == 32, "EdgeAoS must be 32 bytes");

struct VertAoS {
    double x, y, z, angle;
    int32_t firstEdge;
    int32_t flags;
};
static_assert(sizeof(VertAoS) == 40, "VertAoS must be 40 bytes");

struct HostMesh {
    int nv = 0, nf = 0, ne = 0;
    std::vector<double> xyz;          // 3*nv
    std::vector<int32_t> tri;         // 3*nf
    std::vector<int32_t> twin;        // ne
    std::vector<double> len;          // ne
    std::vector<double> angle;        // nv
    std::vector<int32_t> firstEdge;   // nv
    std::vector<uint8_t> flags;       // nv
    std::vector<double> cornerA;      // ne: corner angle at verts[0](e) of face(e)
    std::vector<double> cornerB;      // ne: corner angle at verts[1](e) of face(e)
};

// Builds the half-edge structure; returns an empty string or an error text.
std::string build_host_mesh(const double* xyz, int nv, const int32_t* tri, int nf, HostMesh& m);
std::string load_obj(const char* path, std::vector<double>& xyz, std::vector<int32_t>& tri);
// Polyline of the SVG arc u -> v from its two stored key points (geodesic.cpp):
// appends the crossings' xyz and v's (not u's) to pts; dstXTraced (optional)
// receives the traced position on the last key point's edge.
std::string trace_arc(const HostMesh& m, int u, int v, int srcEdge, double srcX, int dstEdge, int hopCap,
                      std::vector<double>& pts, double* dstXTraced);

// Device images of one mesh (all device pointers).
struct DevMesh {
    int nv = 0, nf = 0, ne = 0;
    FaceRec* face = nullptr;        // nf
    double* cornerA = nullptr;      // ne
    double* cornerB = nullptr;      // ne
    int32_t* firstEdge = nullptr;   // nv
    uint8_t* flags = nullptr;       // nv
    double* pos = nullptr;          // 4*nv (x, y, z, 0) for the A* heuristic
    EdgeAoS* edges = nullptr;       // ne, reference layout (plain kernel)
    FanRec* fan = nullptr;          // ne: the fans of all vertices, vertex by vertex (null: not manifold)
    int32_t* fanIdx = nullptr;      // ne: position of half-edge e in the fan of verts[0](e)
    int32_t* fanBase = nullptr;     // nv + 1: first fan position of each vertex
    VertAoS* verts = nullptr;       // nv, reference layout (plain kernel)
};

}  // namespace svg
