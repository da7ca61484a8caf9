// dambreak_render.cpp - the DamBreak demo (DamBreak.Demo/main.cpp) without a
// window: the demo's system (a 128 x 128 column, one wall layer, grid
// 512 x 256 x 4, r = 1/256), its colour buffer and its camera (translate
// (0.5, 0, -2.5), 60 degree field of view, white background), stepped and
// drawn with the headless ParticleRenderer; every frame is written as a
// binary PPM, frame<k>.ppm.
//
//   dambreak_render [output_dir] [frames] [steps_per_frame] [width] [height]
//   (defaults: ".", 3, 100, and the demo's 1280 x 1024 window)
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "cmag/render.hpp"
#include "cmag/systems.hpp"

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    const int frames = argc > 2 ? std::atoi(argv[2]) : 3;
    const int steps = argc > 3 ? std::atoi(argv[3]) : 100;
    const int width = argc > 4 ? std::atoi(argv[4]) : 1280;
    const int height = argc > 5 ? std::atoi(argv[5]) : 1024;
    if (frames < 1 || steps < 0 || width < 1 || height < 1) {
        std::fprintf(stderr, "usage: dambreak_render [dir] [frames >= 1] [steps >= 0] [width] [height]\n");
        return 2;
    }
    try {
        const unsigned num = 128;
        DamBreakSystem psystem(make_uint3(num, num, 1), 1, make_uint3(4 * num, 2 * num, 4), 1.0f / (2 * num),
                               true);
        psystem.reset();
        ParticleRenderer renderer;
        renderer.setWindowSize(width, height);
        renderer.setFOV(60.0f);
        renderer.setradius(psystem.getParticleRadius());
        renderer.setColorBuffer(psystem.getCudaColorVBO());
        for (int k = 0; k < frames; ++k) {
            for (int s = 0; s < steps; ++s) psystem.update();
            renderer.setVertexBuffer(psystem.getCudaPosVBO(), psystem.getNumParticles());
            renderer.display();
            const std::vector<uint8_t> rgba = renderer.read();
            const std::string file = dir + "/frame" + std::to_string(k) + ".ppm";
            std::ofstream out(file, std::ios::binary);
            if (!out) {
                std::fprintf(stderr, "cannot write %s\n", file.c_str());
                return 1;
            }
            out << "P6\n" << width << " " << height << "\n255\n";
            std::vector<uint8_t> rgb((size_t)width * height * 3);
            for (size_t p = 0; p < (size_t)width * height; ++p)
                for (int c = 0; c < 3; ++c) rgb[3 * p + c] = rgba[4 * p + c];
            out.write(reinterpret_cast<const char*>(rgb.data()), (std::streamsize)rgb.size());
            std::printf("frame=%d t=%.6g file=%s\n", k, psystem.getElapsedTime(), file.c_str());
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
