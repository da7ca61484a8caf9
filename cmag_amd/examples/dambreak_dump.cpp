// dambreak_dump.cpp - the dam-break particle snapshot dump on the C++ interface.
//
// Same scenario and file format as DamBreak.Report/dump.h:16-87: a 128 x 128
// water column (one wall layer, grid 512 x 256 x 4, r = 1/256), relaxed behind
// the right wall, which is then moved twice by the column width
// (changeRightBoundary), and stepped until each snapshot time.  Each snapshot
// dump<t>.dat (t as the stream prints it, its second character replaced by
// 'x': 1 -> dump1x.dat, 1.3 -> dump1x3.dat) starts with the header line
// "x y z w density pressure " and lists the fluid particles in sorted order
// (index[i] for i = 0, 1, ...) as "x y z w vx vy": the reference labels the
// last two columns density and pressure but writes the x and y velocity, and
// so does this program.  The device arrays are copied from the borrowed device
// pointers, as the reference copies them with thrust::copy.
//
//   dambreak_dump [output_dir] [relax_seconds] [t1 t2 ...]
//   (defaults: ".", 1.0 and the reference's times 1.0 1.3 1.7 2.3 2.5 2.7 3.2 3.7)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "cmag/systems.hpp"

namespace {

void copy_down(std::vector<char>& host, const void* dev, size_t bytes) {
    host.resize(bytes);
    if (hipMemcpy(host.data(), dev, bytes, hipMemcpyDeviceToHost) != hipSuccess)
        throw std::runtime_error("hipMemcpy from a device array failed");
}

}  // namespace

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    const float relax = argc > 2 ? std::strtof(argv[2], nullptr) : 1.0f;
    std::vector<float> frames;
    for (int k = 3; k < argc; ++k) frames.push_back(std::strtof(argv[k], nullptr));
    if (frames.empty()) frames = {1.0f, 1.3f, 1.7f, 2.3f, 2.5f, 2.7f, 3.2f, 3.7f};
    const int num = 128;
    try {
        DamBreakSystem psystem(make_uint3(num, num, 1), 1, make_uint3(4 * num, 2 * num, 4), 1.0f / (2 * num),
                               false);
        psystem.reset();
        while (psystem.getElapsedTime() < relax) psystem.update();
        psystem.changeRightBoundary();
        psystem.changeRightBoundary();
        const int n = psystem.getNumParticles();
        std::vector<char> pos, idx, vel;
        int steps = 0;
        for (float t : frames) {
            while (psystem.getElapsedTime() < t) {
                psystem.update();
                ++steps;
            }
            copy_down(pos, psystem.getCudaPosVBO(), sizeof(float4) * (size_t)n);
            copy_down(idx, psystem.getCudaIndex(), sizeof(unsigned int) * (size_t)n);
            copy_down(vel, psystem.getCudaVelVBO(), sizeof(float4) * (size_t)n);
            const float4* p = reinterpret_cast<const float4*>(pos.data());
            const unsigned int* ix = reinterpret_cast<const unsigned int*>(idx.data());
            const float4* v = reinterpret_cast<const float4*>(vel.data());
            std::ostringstream name;
            name << t;
            const std::string file = dir + "/dump" + name.str().replace(1, 1, "x") + ".dat";
            std::ofstream fp(file);
            if (!fp) {
                std::fprintf(stderr, "cannot write %s\n", file.c_str());
                return 1;
            }
            fp << "x " << "y " << "z " << "w " << "density " << "pressure " << std::endl;
            int rows = 0;
            for (int i = 0; i < n; ++i) {
                const unsigned int k = ix[i];
                if (k >= (unsigned int)n || p[k].w != CMAG_DB_FLUID) continue;
                fp << p[k].x << " " << p[k].y << " " << p[k].z << " " << p[k].w << " " << v[k].x << " " << v[k].y
                   << " " << std::endl;
                ++rows;
            }
            std::printf("t=%g elapsed=%.7g steps=%d file=%s rows=%d\n", t, psystem.getElapsedTime(), steps,
                        file.c_str(), rows);
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
