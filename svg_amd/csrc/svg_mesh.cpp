// Host half-edge builder (Mesh.cu:24-126 semantics; on open meshes boundary
// twins are -1, a boundary vertex starts its fan at the boundary and is never a
// saddle).
// Compiled with -ffp-contract=off so lengths and angle sums are the same
// correctly rounded values on every build.
#include "svg_mesh.h"
#include "host_parallel.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <system_error>
#include <thread>
#include <vector>

namespace svg {

static const double kPi = 3.141592653589793238;   // Geometry.cuh:18


static double clamp_acos(double c)
{
    if (c > 1.0) c = 1.0; else if (c < -1.0) c = -1.0;
    return std::acos(c);
}

std::string build_host_mesh(const double* xyz, int nv, const int32_t* tri, int nf, HostMesh& m)
{
    if (nv <= 0 || nf <= 0) return "empty mesh";
    if ((long long)nf * 3 > 0x7fffffffLL) return "too many faces";
    for (long long i = 0; i < 3LL * nf; ++i)
        if (tri[i] < 0 || tri[i] >= nv) return "triangle index out of range";
    for (int f = 0; f < nf; ++f) {
        const int32_t a = tri[3 * (size_t)f], b = tri[3 * (size_t)f + 1], c = tri[3 * (size_t)f + 2];
        if (a == b || b == c || a == c) return "face " + std::to_string(f) + " repeats a vertex";
    }
    for (int v = 0; v < nv; ++v)
        if (!std::isfinite(xyz[3 * (size_t)v]) || !std::isfinite(xyz[3 * (size_t)v + 1]) ||
            !std::isfinite(xyz[3 * (size_t)v + 2]))
            return "vertex " + std::to_string(v) + " has a non-finite coordinate";
    m.nv = nv; m.nf = nf; m.ne = 3 * nf;
    m.xyz.assign(xyz, xyz + 3 * (size_t)nv);
    m.tri.assign(tri, tri + 3 * (size_t)nf);
    m.twin.assign(m.ne, -1);                                   // boundary: -1
    m.len.resize(m.ne);
    auto next = [](int e) { return 3 * (e / 3) + (e % 3 + 1) % 3; };
    auto prev = [](int e) { return 3 * (e / 3) + (e % 3 + 2) % 3; };
    auto v0 = [&](int e) { return tri[e]; };
    parallel_for(nf, [&](int f0, int f1) {
        for (int f = f0; f < f1; ++f) {
            for (int j = 0; j < 3; ++j) {
                int a = tri[3 * f + j], b = tri[3 * f + (j + 1) % 3];
                double x = xyz[3 * a] - xyz[3 * b];
                double y = xyz[3 * a + 1] - xyz[3 * b + 1];
                double z = xyz[3 * a + 2] - xyz[3 * b + 2];
                m.len[3 * f + j] = std::sqrt(x * x + y * y + z * z);
            }
        }
    });
    for (int e = 0; e < m.ne; ++e)                   // coincident vertices would put 0/0 into the window geometry
        if (!(m.len[e] > 0.0)) return "face " + std::to_string(e / 3) + " has an edge of length 0";
    // Twins: the half-edge a -> b pairs with b -> a (Mesh.cu:79-89 pairs the
    // half-edges of one unordered vertex pair).  A direction met twice (an
    // edge of more than two faces, or faces of opposite orientation) is
    // refused: the reference would overwrite or mis-pair the twins, and
    // every walk around a vertex relies on twin(twin(e)) = e running the
    // other way.  The pairs are found in each vertex's list of outgoing
    // half-edges (destinations stored alongside).
    {
        std::vector<int32_t> ofs((size_t)nv + 1, 0), outE((size_t)m.ne), outD((size_t)m.ne);
        std::vector<uint8_t> bad((size_t)m.ne, 0);
        for (int e = 0; e < m.ne; ++e) ofs[tri[e] + 1]++;
        for (int v = 0; v < nv; ++v) ofs[v + 1] += ofs[v];
        {
            std::vector<int32_t> fill(ofs.begin(), ofs.end() - 1);
            for (int e = 0; e < m.ne; ++e) {
                const int k = fill[tri[e]]++;
                outE[k] = e;
                outD[k] = tri[next(e)];
            }
        }
        parallel_for(m.ne, [&](int e0, int e1) {
            for (int e = e0; e < e1; ++e) {
                const int32_t a = tri[e], b = tri[next(e)];
                int same = 0, nopp = 0, opp = -1;
                for (int k = ofs[a]; k < ofs[a + 1]; ++k) same += outD[k] == b;        // e itself included
                for (int k = ofs[b]; k < ofs[b + 1]; ++k)
                    if (outD[k] == a) { opp = outE[k]; ++nopp; }
                if (same > 1 || nopp > 1) bad[e] = 1;
                else m.twin[e] = opp;
            }
        });
        for (int e = 0; e < m.ne; ++e) {
            if (!bad[e]) continue;
            const int32_t a = tri[e], b = tri[next(e)];
            const std::string edge = "(" + std::to_string(std::min(a, b)) + ", " + std::to_string(std::max(a, b)) + ")";
            int same = 0, nopp = 0, other = -1;
            for (int k = ofs[a]; k < ofs[a + 1]; ++k)
                if (outD[k] == b) { ++same; if (outE[k] != e) other = outE[k]; }
            for (int k = ofs[b]; k < ofs[b + 1]; ++k) nopp += outD[k] == a;
            if (same + nopp > 2) return "edge " + edge + " is shared by more than two faces";
            return "faces " + std::to_string(std::min(e, other) / 3) + " and " + std::to_string(std::max(e, other) / 3) +
                   " are oriented oppositely across edge " + edge;
        }
    }
    auto v1 = [&](int e) { return tri[next(e)]; };
    // Corner angles used by the pseudo-source fan sweeps, with the operand
    // order of ICHDevice.cu:830 / 847 (and 876 / 893):
    //   A(e): L0 = len(e), L1 = len(prev e), L2 = len(next e)   (left sweep)
    //   B(e): L0 = len(e), L1 = len(next e), L2 = len(prev e)   (right sweep)
    m.cornerA.resize(m.ne);
    m.cornerB.resize(m.ne);
    parallel_for(m.ne, [&](int e0, int e1) {
        for (int e = e0; e < e1; ++e) {
            double L0 = m.len[e];
            double La1 = m.len[prev(e)], La2 = m.len[next(e)];
            m.cornerA[e] = clamp_acos((L0 * L0 + La1 * La1 - La2 * La2) / (2.0 * L0 * La1));
            double Lb1 = m.len[next(e)], Lb2 = m.len[prev(e)];
            m.cornerB[e] = clamp_acos((L0 * L0 + Lb1 * Lb1 - Lb2 * Lb2) / (2.0 * L0 * Lb1));
        }
    });
    // Vertex angle sums (Mesh.cu:107-121) add, in half-edge order, the corner
    // at verts[0]: (l0^2 + l2^2 - l1^2) / (2 l0 l2) with l1 = len(next e),
    // l2 = len(prev e) is A(e)'s expression operand for operand, so the
    // stored A(e) is that corner to the bit.
    m.angle.assign(nv, 0.0);
    m.firstEdge.assign(nv, -1);
    for (int e = 0; e < m.ne; ++e) {
        m.angle[v0(e)] += m.cornerA[e];
        m.firstEdge[v0(e)] = e;
    }
    std::vector<uint8_t> bnd(nv, 0);
    for (int e = 0; e < m.ne; ++e)
        if (m.twin[e] == -1) { bnd[v0(e)] = 1; bnd[v1(e)] = 1; }
    for (int e = 0; e < m.ne; ++e)                             // boundary fan start
        if (bnd[v0(e)] && m.twin[prev(e)] == -1) m.firstEdge[v0(e)] = e;
    m.flags.assign(nv, 0);
    for (int v = 0; v < nv; ++v) {
        uint8_t f = 0;
        if (bnd[v] || m.angle[v] < 2.0 * kPi) f |= VF_NOT_SADDLE;     // boundary: never a saddle
        if (!bnd[v] && m.angle[v] > 2.0 * kPi) f |= VF_SADDLE;
        if (bnd[v]) f |= VF_BOUNDARY;
        m.flags[v] = f;
    }
    return std::string();
}

// OBJ reader.  The reference (Mesh.cu:41-60) reads "v x y z" lines and the
// first three plain 1-based indices of "f" lines, skipping everything else;
// files of that form load identically here.  Beyond it, as real meshes need:
// a face corner may be written "i", "i/t", "i//n" or "i/t/n" (only the vertex
// index is used), a negative index counts back from the last vertex read, a
// face with more than three corners is split into the fan (c0, ci, ci+1) (the
// reference would keep its first triangle only), and "#" starts a comment.
// Texture and normal lines ("vt", "vn") are skipped, as there.
std::string load_obj(const char* path, std::vector<double>& xyz, std::vector<int32_t>& tri)
{
    std::ifstream in(path);
    if (!in) return std::string("cannot open file ") + path;
    std::string line;
    xyz.clear(); tri.clear();
    long lineNo = 0;
    std::vector<long> corner;
    while (std::getline(in, line)) {
        ++lineNo;
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        if (line.size() < 2) continue;
        std::istringstream ss(line);
        std::string tag;
        ss >> tag;
        if (tag == "v") {
            double x, y, z;
            ss >> x >> y >> z;
            if (!ss) return "line " + std::to_string(lineNo) + ": bad vertex line: " + line;
            xyz.push_back(x); xyz.push_back(y); xyz.push_back(z);
        } else if (tag == "f") {
            corner.clear();
            std::string tok;
            const long nv = (long)(xyz.size() / 3);
            while (ss >> tok) {
                const size_t slash = tok.find('/');
                const std::string head = tok.substr(0, slash);
                char* end = nullptr;
                const long idx = std::strtol(head.c_str(), &end, 10);
                if (head.empty() || *end != '\0' || idx == 0)
                    return "line " + std::to_string(lineNo) + ": bad face corner '" + tok + "'";
                corner.push_back(idx > 0 ? idx - 1 : nv + idx);
            }
            if (corner.size() < 3) return "line " + std::to_string(lineNo) + ": face with fewer than 3 corners";
            for (size_t i = 1; i + 1 < corner.size(); ++i) {
                tri.push_back((int32_t)corner[0]);
                tri.push_back((int32_t)corner[i]);
                tri.push_back((int32_t)corner[i + 1]);
            }
        }
    }
    return std::string();
}

}  // namespace svg
