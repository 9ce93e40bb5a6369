// Handle layouts shared by the C ABI translation units.
#pragma once

#include <exception>
#include <mutex>
#include <new>
#include <string>

#include "svg_amd.h"
#include "svg_build.h"
#include "svg_graph.h"
#include "svg_mesh.h"

namespace svg { struct SolverState; struct PointWork; }   // solve.h, point.h

struct svg_mesh_s {
    svg::HostMesh h;
    svg::DevMesh d;
    std::mutex mu;                        // autotuner records
    std::mutex upMu;                      // device upload of the mesh arrays
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

// Every entry point of the C ABI runs its body through this guard, so no
// C++ exception reaches a C caller: a failed host allocation is SVG_E_MEM,
// anything else thrown is SVG_E_INTERNAL with its message.
template <class R, class F>
R svg_guarded(F&& body) noexcept
{
    try {
        return body();
    } catch (const std::bad_alloc&) {
        return (R)svg_set_error(SVG_E_MEM, "host allocation failed");
    } catch (const std::exception& e) {
        return (R)svg_set_error(SVG_E_INTERNAL, std::string("host failure: ") + e.what());
    } catch (...) {
        return (R)svg_set_error(SVG_E_INTERNAL, "host failure");
    }
}
int svg_ensure_device();
void svg_solver_release(svg_graph_s* g);
void svg_point_release(svg_graph_s* g);
int svg_mesh_upload_impl(svg_mesh_s* m, bool aos);
