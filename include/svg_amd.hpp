// svg_amd.hpp -- header-only C++ mirror of the reference Mesh / SVG classes
// (Mesh.cuh:32-52, SVG.cuh:49-82) over the C ABI of svg_amd.h.  A caller of
// the reference keeps its call sequence:
//
//   Mesh mesh;  mesh.LoadFromFile("in.obj");  mesh.copyToGPU();
//   SVG svg;    svg.AssignMesh(&mesh);  svg.SetParameters(K, 10, 2);
//   svg.Allocation();  svg.ConstructSVG();  svg.Free();
//   svg.CopySVGToHost();  svg.SaveSVGToFile("in.svg");
//   double d = svg.SolveSSSD(s, t);
//
// Errors throw svg_amd::Error (the reference printed and exited).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "svg_amd.h"

namespace svg_amd {

struct Error : std::runtime_error {
    explicit Error(const std::string& m) : std::runtime_error(m) {}
};

inline void check(int rc)
{
    if (rc < 0) throw Error(std::string("svg_amd: ") + svg_last_error());
}

class Mesh {
public:
    Mesh() = default;
    Mesh(const Mesh&) = delete;
    Mesh& operator=(const Mesh&) = delete;
    ~Mesh() { clear(); }

    bool LoadFromFile(const char* fileName)          // Mesh.cu:24-126
    {
        clear();
        return svg_mesh_load_obj(fileName, &h_) == SVG_OK;
    }
    void FromArrays(const double* xyz, int nv, const int32_t* tri, int nf)
    {
        clear();
        check(svg_mesh_from_arrays(xyz, nv, tri, nf, &h_));
    }
    bool copyToGPU() { check(svg_mesh_upload(h_)); return true; }   // Mesh.cu:128-142
    void clear() { if (h_) { svg_mesh_free(h_); h_ = nullptr; } }    // Mesh.cu:144-159
    void clearGPU() { clear(); }
    int vertNum() const { int n = 0; svg_mesh_info(h_, &n, nullptr, nullptr); return n; }
    int faceNum() const { int n = 0; svg_mesh_info(h_, nullptr, &n, nullptr); return n; }
    int edgeNum() const { int n = 0; svg_mesh_info(h_, nullptr, nullptr, &n); return n; }
    svg_mesh_t handle() const { return h_; }

private:
    svg_mesh_t h_ = nullptr;
};

class SVG {
public:
    struct SVGNode {                                  // SVG.cuh:21-27
        int adjNode;
        double geodDist;
        unsigned nextToSrcEdge, nextToDstEdge;
        double nextToSrcX, nextToDstX;
    };
    struct GraphDistInfo {                            // SVG.cuh:29-37
        double dist;
        int pathParentIndex;
        int srcId;
    };

    SVG() = default;
    SVG(const SVG&) = delete;
    SVG& operator=(const SVG&) = delete;
    ~SVG() { FreeSVGStructure(); }

    void AssignMesh(Mesh* mesh, Mesh* = nullptr) { mesh_ = mesh; }      // SVG.cu:109-112
    void SetParameters(int K, unsigned splitInfoCoef, unsigned vertInfoCoef)
    {
        svg_build_params_default(&p_, K);
        p_.split_coef = splitInfoCoef;
        p_.vert_coef = vertInfoCoef;
    }
    bool Allocation() { return true; }                 // scratch is sized inside the build
    void SetMode(int mode) { p_.mode = mode; }         // SVG_MODE_BATCHED (default) / _PLAIN
    svg_build_params& Params() { return p_; }
    void ConstructSVG()                                // SVG.cu:160-182
    {
        FreeSVGStructure();
        check(svg_build(mesh_->handle(), &p_, &g_));
    }
    void Free() {}                                     // build scratch is already released
    void FreeSVGStructure()                            // SVG.cu:152-158
    {
        if (g_) { svg_graph_free(g_); g_ = nullptr; }
        tails_.clear(); rows_.clear();
    }
    void CopySVGToHost()                               // SVG.cu:184-206
    {
        int V = 0, K = 0;
        long long arcs = 0;
        check(svg_graph_info(g_, &V, &K, &arcs));
        const size_t n = (size_t)V * (size_t)K;
        std::vector<int32_t> adj(n), se(n), de(n);
        std::vector<double> w(n), sx(n), dx(n);
        tails_.resize((size_t)V);
        check(svg_graph_to_host(g_, tails_.data(), adj.data(), w.data(), se.data(), de.data(), sx.data(),
                                dx.data()));
        rows_.resize(n);
        for (size_t i = 0; i < n; ++i)
            rows_[i] = SVGNode{adj[i], w[i], (unsigned)se[i], (unsigned)de[i], sx[i], dx[i]};
        K_ = K;
    }
    const SVGNode* Row(int v) const { return rows_.data() + (size_t)v * (size_t)K_; }
    int Tail(int v) const { return tails_[(size_t)v]; }
    void SaveSVGToFile(const char* fileName) { check(svg_graph_save(g_, fileName)); }     // SVG.cu:208-217
    void LoadSVGFromFile(const char* fileName)                                            // SVG.cu:219-239
    {
        FreeSVGStructure();
        check(svg_graph_load(mesh_->handle(), p_.K, fileName, &g_));
    }
    double SolveSSSD(int s, int t)                     // SVG.cu:241-251
    {
        double d = 0.0;
        check(svg_sssd(g_, s, t, &d, nullptr, 0, nullptr));
        return d;
    }
    // SVG.cu:253-352: distance between point p0 of face f0 and point p1 of
    // face f1 (xyz[3] each); res receives geodDist and the key points,
    // *saddleVertNearDst the last graph vertex (-1: answered directly).
    void SolveSSSD(int f0, const double* p0, int f1, const double* p1, SVGNode* res, int* saddleVertNearDst)
    {
        double d = 0.0, kx[2] = {0.0, 0.0};
        int32_t ke[2] = {-1, -1}, last = -1;
        int route = 0;
        check(svg_sssd_point(g_, f0, p0, f1, p1, &d, ke, kx, &last, &route, nullptr, 0, nullptr));
        res->adjNode = -1;
        res->geodDist = d;
        res->nextToSrcEdge = (unsigned)ke[0]; res->nextToDstEdge = (unsigned)ke[1];
        res->nextToSrcX = kx[0]; res->nextToDstX = kx[1];
        if (saddleVertNearDst) *saddleVertNearDst = last;
    }
    // Many point pairs at once: faces[2n] = (f0, f1), points[6n] = (p0, p1).
    // route[q] (optional) is 0 direct, 1 through the graph, -1 nothing joins
    // the points, -2 a point lies on a vertex of its face (NaN distance).
    void SolveSSSDBatch(int n, const int32_t* faces, const double* points, std::vector<SVGNode>& res,
                        std::vector<int>* saddleVertNearDst = nullptr, std::vector<int>* route = nullptr)
    {
        std::vector<double> d((size_t)n), kx(2 * (size_t)n);
        std::vector<int32_t> ke(2 * (size_t)n), last((size_t)n), rt((size_t)n);
        check(svg_sssd_point_batch(g_, n, faces, points, d.data(), ke.data(), kx.data(), last.data(), rt.data()));
        res.resize((size_t)n);
        for (int q = 0; q < n; ++q)
            res[(size_t)q] = SVGNode{-1, d[(size_t)q], (unsigned)ke[2 * (size_t)q], (unsigned)ke[2 * (size_t)q + 1],
                                     kx[2 * (size_t)q], kx[2 * (size_t)q + 1]};
        if (saddleVertNearDst) saddleVertNearDst->assign(last.begin(), last.end());
        if (route) route->assign(rt.begin(), rt.end());
    }
    void SolveMSAD(const int* sources, const double* weights, int Ns, std::vector<GraphDistInfo>& out)
    {                                                  // SVG.cu:354-367
        int V = 0;
        check(svg_graph_info(g_, &V, nullptr, nullptr));
        std::vector<double> d((size_t)V);
        std::vector<int32_t> sid((size_t)V), par((size_t)V);
        check(svg_msad(g_, sources, weights, Ns, d.data(), sid.data(), par.data()));
        out.resize((size_t)V);
        for (int v = 0; v < V; ++v) out[(size_t)v] = GraphDistInfo{d[(size_t)v], par[(size_t)v], sid[(size_t)v]};
    }
    svg_graph_t handle() const { return g_; }

private:
    Mesh* mesh_ = nullptr;
    svg_build_params p_{};
    svg_graph_t g_ = nullptr;
    std::vector<int32_t> tails_;
    std::vector<SVGNode> rows_;
    int K_ = 0;
};

}  // namespace svg_amd
