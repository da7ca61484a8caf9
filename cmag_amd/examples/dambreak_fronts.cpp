// dambreak_fronts.cpp - the dam-break surge-front and column-height trackers on
// the C++ interface.
//
// Same scenario as DamBreak.Report/XFrontTest.h:16-90 and YFrontTest.h:16-93:
// a 128 x 256 water column (r = 1/256, one wall layer, grid 512 x 256 x 4),
// relaxed behind the right wall for 1.5 s, released (removeRightBoundary,
// which restarts the clock), and sampled at the dimensionless times of Martin
// & Moyce (1952, Tables 2 and 6 for n^2 = 2):
//   XFrontOutput  T* elapsed Z*_exp Z*   with Z = (front + r - origin.x)/w,
//                 w = (fx + offset) 2r, T* = t sqrt(2|g|/w)
//   YFrontOutput  T* elapsed H*_exp H*   with H = (top + r - origin.y)/H0,
//                 H0 = the relaxed column height, T* = t sqrt(2|g|/H0)
// One run serves both trackers (the reference relaxes the same system twice).
// The device arrays are read back with getArray instead of thrust::copy.
//
//   dambreak_fronts [output_dir]
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "cmag/systems.hpp"

namespace {

// Martin & Moyce: (dimensionless time, surge front / column height)
const float kSurge[][2] = {{0.00f, 1.00f}, {0.41f, 1.11f}, {0.84f, 1.22f}, {1.19f, 1.44f}, {1.43f, 1.67f},
                           {1.63f, 1.89f}, {1.83f, 2.11f}, {1.98f, 2.33f}, {2.20f, 2.56f}, {2.32f, 2.78f},
                           {2.51f, 3.00f}, {2.65f, 3.22f}, {2.83f, 3.44f}, {2.97f, 3.67f}, {3.11f, 3.89f}};
const float kHeight[][2] = {{0.00f, 1.00f}, {0.56f, 0.94f}, {0.77f, 0.89f}, {0.93f, 0.83f}, {1.08f, 0.78f},
                            {1.28f, 0.72f}, {1.46f, 0.67f}, {1.66f, 0.61f}, {1.84f, 0.56f}, {2.00f, 0.50f},
                            {2.21f, 0.44f}, {2.45f, 0.39f}, {2.70f, 0.33f}, {3.06f, 0.28f}, {3.44f, 0.22f}};
constexpr int kFrames = 15;

// the largest coordinate a (0: x, 1: y) of the fluid particles
float fluid_max(const std::vector<float>& pos, int n, int a) {
    float m = -INFINITY;
    for (int i = 0; i < n; ++i)
        if (pos[4 * (size_t)i + 3] == CMAG_DB_FLUID) m = std::fmax(m, pos[4 * (size_t)i + a]);
    return m;
}

}  // namespace

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    const int num = 128;
    const uint3 fluid = make_uint3(num, 2 * num, 1);
    const uint3 gridSize = make_uint3(512, 256, 4);
    const int boundaryOffset = 1;
    const float radius = 1.0f / (2 * num);
    try {
        DamBreakSystem sys(fluid, boundaryOffset, gridSize, radius, false);
        sys.reset();
        const int n = sys.getNumParticles();
        std::vector<float> pos(4 * (size_t)n);
        while (sys.getElapsedTime() < 1.5f) sys.update();
        sys.removeRightBoundary();
        std::printf("relaxed: particles=%d\n", n);

        sys.getArray(DamBreakSystem::POSITION, pos.data(), 0, n);
        const float g = std::fabs(sys.getGravity().y);
        const float width = (fluid.x + boundaryOffset) * 2 * radius;
        const float height = fluid_max(pos, n, 1) + radius - sys.getWorldOrigin().y;
        const float scale_x = std::sqrt(2 * g / width), scale_y = std::sqrt(2 * g / height);

        FILE* fx = std::fopen((dir + "/XFrontOutput").c_str(), "w");
        FILE* fy = std::fopen((dir + "/YFrontOutput").c_str(), "w");
        if (!fx || !fy) {
            std::fprintf(stderr, "cannot write to %s\n", dir.c_str());
            return 1;
        }
        int kx = 0, ky = 0, steps = 0;
        while (kx < kFrames || ky < kFrames) {
            const bool next_x = kx < kFrames && (ky >= kFrames || kSurge[kx][0] / scale_x <= kHeight[ky][0] / scale_y);
            const float T = next_x ? kSurge[kx][0] : kHeight[ky][0];
            const float scale = next_x ? scale_x : scale_y;
            while (sys.getElapsedTime() * scale < T) {
                sys.update();
                ++steps;
            }
            sys.getArray(DamBreakSystem::POSITION, pos.data(), 0, n);
            if (next_x) {
                const float z = (fluid_max(pos, n, 0) + radius - sys.getWorldOrigin().x) / width;
                std::fprintf(fx, "%f %f %f %f \n", T, sys.getElapsedTime(), kSurge[kx][1], z);
                ++kx;
            } else {
                const float h = (fluid_max(pos, n, 1) + radius - sys.getWorldOrigin().y) / height;
                std::fprintf(fy, "%f %f %f %f \n", T, sys.getElapsedTime(), kHeight[ky][1], h);
                ++ky;
            }
        }
        std::fclose(fx);
        std::fclose(fy);
        std::printf("released: steps=%d width=%g height=%g\n", steps, width, height);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
