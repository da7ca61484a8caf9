// poiseuille_report.cpp - the Poiseuille report run on the C++ interface.
//
// Same scenario as Poiseuille.Report/dump.h:15-86: a 16 x 58 fluid channel
// between three wall layers (grid 16 x 64 x 4, r = 1/(2*58*1000)), stepped
// until each report time, then the x velocity against the height above the
// bottom wall of the fluid particles in the column 0 < x < 2r is written to
// XVelocityYPosition<t>.dat (first line "0.0 0.0", last "0.000000 0.001000",
// the format the reference's Maple script XVelocityYPosition.m reads).
// The device arrays are read back with getArray instead of thrust::copy.
//
//   poiseuille_report [output_dir]
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "cmag/systems.hpp"

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    const int boundaryOffset = 3;
    const uint3 gridSize = make_uint3(16, 64, 4);
    const float radius = 1.0f / (2 * (gridSize.y - 2 * boundaryOffset) * 1000);
    const uint3 fluid = make_uint3(gridSize.x, gridSize.y - 2 * boundaryOffset, 1);
    try {
        PoiseuilleFlowSystem psystem(fluid, boundaryOffset, gridSize, radius, false);
        psystem.reset();
        const int n = psystem.getNumParticles();
        std::vector<float> pos(4 * (size_t)n), vel(4 * (size_t)n);
        const float slices[] = {0.0225f, 0.045f, 0.1125f, 0.225f, 1.0f};
        int steps = 0;
        for (float t : slices) {
            while (psystem.getElapsedTime() < t) {
                psystem.update();
                ++steps;
            }
            psystem.getArray(PoiseuilleFlowSystem::POSITION, pos.data(), 0, n);
            psystem.getArray(PoiseuilleFlowSystem::VELOCITY, vel.data(), 0, n);
            char name[256];
            std::snprintf(name, sizeof(name), "%s/XVelocityYPosition%g.dat", dir.c_str(), t);
            FILE* f = std::fopen(name, "w");
            if (!f) {
                std::fprintf(stderr, "cannot write %s\n", name);
                return 1;
            }
            std::fprintf(f, "0.0 0.0\n");
            int written = 0;
            for (int i = 0; i < n; ++i) {
                const float x = pos[4 * i], y = pos[4 * i + 1], w = pos[4 * i + 3];
                if (x > 0 && x < 2 * radius && w == 0.0f) {
                    std::fprintf(f, "%.9g %.9g\n", vel[4 * i],
                                 y + std::fabs(psystem.getWorldOrigin().y) - boundaryOffset * 2 * radius);
                    ++written;
                }
            }
            std::fprintf(f, "0.000000 0.001000\n");
            std::fclose(f);
            std::printf("t=%g elapsed=%.7g steps=%d particles=%d\n", t, psystem.getElapsedTime(), steps, written);
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
