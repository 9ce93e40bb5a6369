// Handle layouts shared by the C ABI translation units.
#pragma once

#include <mutex>
#include <string>

#include "svg_amd.h"
#include "svg_build.h"
#include "svg_graph.h"
#include "svg_mesh.h"

namespace svg { struct SolverState; struct PointWork; }   // solve.h, point.h

struct svg_mesh_s {
    svg::HostMesh h;
    svg::DevMesh d;
    std::mutex mu;
    std::vector<svg::TuneRecord> tunes;   // batched-build autotuner state, per (K, coefficients)
};

struct svg_graph_s {
    const svg_mesh_s* mesh = nullptr;
    svg::DevGraph d;
    svg::BuildStats stats;
    long long arcs = 0;
    svg::SolverState* solver = nullptr;   // lazily created query workspace
    svg::PointWork* point = nullptr;      // lazily created surface-point query buffers
    std::mutex mu;
};

int svg_set_error(int code, const std::string& msg);
int svg_ensure_device();
void svg_solver_release(svg_graph_s* g);
void svg_point_release(svg_graph_s* g);
int svg_mesh_upload_impl(svg_mesh_s* m, bool aos);
