// Command-line driver with the reference's flows (kernel.cu:17-105, 107-221,
// 288-402, 404-456): build + save, A* SSSD, surface-point SSSD, MSAD --
// non-interactive.
//
//   svg_cli build <in.obj> <K> [out.svg] [plain]
//   svg_cli sssd  <in.obj> <K> <s> <t>
//   svg_cli point <in.obj> <K> <f0> <x0> <y0> <z0> <f1> <x1> <y1> <z1>
//   svg_cli msad  <in.obj> <K> <src> [src ...]
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "svg_amd.hpp"

int main(int argc, char** argv)
{
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s build|sssd|point|msad <in.obj> <K> ...\n", argv[0]);
        return 2;
    }
    const std::string cmd = argv[1];
    const int K = std::atoi(argv[3]);
    try {
        svg_amd::Mesh mesh;
        if (!mesh.LoadFromFile(argv[2])) {
            std::fprintf(stderr, "cannot load %s: %s\n", argv[2], svg_last_error());
            return 1;
        }
        mesh.copyToGPU();
        svg_amd::SVG svg;
        svg.AssignMesh(&mesh);
        svg.SetParameters(K, 10, 2);
        if (cmd == "build" && argc > 5 && std::strcmp(argv[5], "plain") == 0) svg.SetMode(SVG_MODE_PLAIN);
        svg.Allocation();
        svg.ConstructSVG();
        svg.Free();
        svg_build_stats st;
        svg_amd::check(svg_build_stats_get(svg.handle(), &st));
        std::printf("Time consumed: %.6f s (device), arcs %lld, average degree %.3f\n", st.seconds, st.arcs,
                    (double)st.arcs / mesh.vertNum());
        if (cmd == "build") {
            std::string out = argc > 4 ? argv[4] : std::string(argv[2]);
            if (argc <= 4) out = out.substr(0, out.rfind('.')) + ".svg";
            svg.SaveSVGToFile(out.c_str());
            std::printf("saved %s\n", out.c_str());
        } else if (cmd == "sssd" && argc >= 6) {
            const double d = svg.SolveSSSD(std::atoi(argv[4]), std::atoi(argv[5]));
            std::printf("Distance: %.17g\n", d);
        } else if (cmd == "point" && argc >= 12) {
            // kernel.cu:160-202: faces and points in, distance and key points out
            const double p0[3] = {std::atof(argv[5]), std::atof(argv[6]), std::atof(argv[7])};
            const double p1[3] = {std::atof(argv[9]), std::atof(argv[10]), std::atof(argv[11])};
            svg_amd::SVG::SVGNode res;
            int lastVert = -1;
            svg.SolveSSSD(std::atoi(argv[4]), p0, std::atoi(argv[8]), p1, &res, &lastVert);
            std::printf("Distance: %.17g\n", res.geodDist);
            std::printf("Next to src edge: %d\nNext to src pos: %.17g\n", (int)res.nextToSrcEdge, res.nextToSrcX);
            std::printf("Next to dst edge: %d\nNext to dst pos: %.17g\n", (int)res.nextToDstEdge, res.nextToDstX);
            std::printf("Last graph vertex: %d\n", lastVert);
        } else if (cmd == "msad" && argc >= 5) {
            std::vector<int> src;
            for (int i = 4; i < argc; ++i) src.push_back(std::atoi(argv[i]));
            std::vector<double> w(src.size(), 0.0);
            std::vector<svg_amd::SVG::GraphDistInfo> info;
            svg.SolveMSAD(src.data(), w.data(), (int)src.size(), info);
            double mx = 0.0;
            for (const auto& g : info)
                if (g.dist < 1e300 && g.dist > mx) mx = g.dist;
            std::printf("MSAD over %zu sources: farthest reached vertex at %.17g\n", src.size(), mx);
        } else {
            std::fprintf(stderr, "bad command\n");
            return 2;
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
