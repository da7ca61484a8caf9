// init.cpp - the reference reset() lattices (host), with the reference's float
// operation order so the device sees bit-identical initial positions.
//
//   DamBreakSystem::initFluid/initBoundaryParticles  DamBreak.Core/fluidSystem.cpp:382-523
//   PoiseuilleFlowSystem::initFluid/...              Poiseuille.Core/poiseuilleFlowSystem.cpp:325-383
//   PeristalsisSystem::initFluid/...                 Peristalsis.Core/PeristalsisSystem.cpp:255-317
//   PeristalsisSystem::CalculateMass                 Peristalsis.Core/PeristalsisSystem.cpp:238-253
#include <cmath>
#include <cstdint>
#include <cstring>

#include "cmag_params.h"

namespace cmag {

namespace {
struct Writer {
    float* a;
    uint32_t n;
    void put(uint32_t i, float x, float y, float z, float w) const {
        if (i >= n) return;
        a[4 * (size_t)i] = x;
        a[4 * (size_t)i + 1] = y;
        a[4 * (size_t)i + 2] = z;
        a[4 * (size_t)i + 3] = w;
    }
};
}  // namespace

void init_dambreak(const cmag_params& P, float* pos4, uint32_t n) {
    std::memset(pos4, 0, sizeof(float) * 4 * (size_t)n);
    Writer W{pos4, n};
    const float r = P.particle_radius, spacing = r * 2.0f;
    const int bo = P.boundary_offset;
    const uint32_t fx = P.fluid_size[0], fy = P.fluid_size[1], fz = P.fluid_size[2];
    const float halfX = P.grid_size[0] * r, halfY = P.grid_size[1] * r, halfZ = P.grid_size[2] * r;
    const float ox = P.world_origin[0], oy = P.world_origin[1], oz = P.world_origin[2];
    for (uint32_t z = 0; z < fz; z++)
        for (uint32_t y = 0; y < fy; y++)
            for (uint32_t x = 0; x < fx; x++)
                W.put((z * fy * fx) + y * fx + x, (spacing * x) + r - halfX + bo * 2 * r,
                      (spacing * y) + r - halfY + bo * 2 * r, (spacing * z) + r - halfZ,
                      CMAG_DB_FLUID);
    uint32_t base = fx * fy * fz;
    // six wall blocks in the reference's order: bottom/left/right type 1, then type 2
    struct Block {
        uint32_t sx, sy;
        float x0, y0, w;
    };
    const Block blocks[6] = {
        {P.grid_size[0] - (bo - 1), 1u, ox + (bo - 1) * 2 * r, oy + (bo - 1) * 2 * r, CMAG_DB_FIRST},
        {1u, P.grid_size[1] - bo, ox + (bo - 1) * 2 * r, oy + (bo + 0) * 2 * r, CMAG_DB_FIRST},
        {1u, P.grid_size[1] - bo, ox + (fx + bo) * 2 * r, oy + bo * 2 * r, CMAG_DB_RIGHT_FIRST},
        {P.grid_size[0], (uint32_t)(bo - 1), ox, oy, CMAG_DB_SECOND},
        {(uint32_t)(bo - 1), P.grid_size[1] - (bo - 1), ox, oy + (bo - 1) * 2 * r, CMAG_DB_SECOND},
        {(uint32_t)(bo - 1), P.grid_size[1] - bo, ox + (fx + bo + 1) * 2 * r, oy + (bo) * 2 * r,
         CMAG_DB_RIGHT_SECOND},
    };
    // each block writes x = (spacing*x + r + origin) + offset, kept in that order
    const float xoff[6] = {(bo - 1) * 2 * r, (bo - 1) * 2 * r, (fx + bo) * 2 * r, 0.0f, 0.0f,
                           (fx + bo + 1) * 2 * r};
    const float yoff[6] = {(bo - 1) * 2 * r, (bo + 0) * 2 * r, bo * 2 * r, 0.0f, (bo - 1) * 2 * r,
                           (bo) * 2 * r};
    const bool has_xoff[6] = {true, true, true, false, false, true};
    const bool has_yoff[6] = {true, true, true, false, true, true};
    for (int b = 0; b < 6; ++b) {
        const Block& B = blocks[b];
        for (uint32_t y = 0; y < B.sy; y++)
            for (uint32_t x = 0; x < B.sx; x++) {
                float px = (spacing * x) + r + ox;
                if (has_xoff[b]) px = px + xoff[b];
                float py = (spacing * y) + r + oy;
                if (has_yoff[b]) py = py + yoff[b];
                W.put(base + y * B.sx + x, px, py, (spacing * 0u) + r + oz, B.w);
            }
        base += B.sx * B.sy;
    }
}

void init_poiseuille(const cmag_params& P, float* pos4, uint32_t n) {
    std::memset(pos4, 0, sizeof(float) * 4 * (size_t)n);
    Writer W{pos4, n};
    const float r = P.particle_radius, spacing = r * 2.0f;
    const int bo = P.boundary_offset;
    const uint32_t fx = P.fluid_size[0], fy = P.fluid_size[1], fz = P.fluid_size[2];
    const float halfX = P.grid_size[0] * r, halfY = P.grid_size[1] * r, halfZ = P.grid_size[2] * r;
    const float ox = P.world_origin[0], oy = P.world_origin[1], oz = P.world_origin[2];
    for (uint32_t z = 0; z < fz; z++)
        for (uint32_t y = 0; y < fy; y++)
            for (uint32_t x = 0; x < fx; x++)
                W.put((z * fy * fx) + y * fx + x, (spacing * x) + r - halfX,
                      (spacing * y) + r - halfY + bo * 2 * r, (spacing * z) + r - halfZ, CMAG_PO_FLUID);
    uint32_t base = fx * fy * fz;
    const uint32_t sx = P.grid_size[0], sy = (uint32_t)bo;
    for (uint32_t y = 0; y < sy; y++)
        for (uint32_t x = 0; x < sx; x++)
            W.put(base + y * sx + x, (spacing * x) + r + ox, (spacing * y) + r + oy,
                  (spacing * 0u) + r + oz, CMAG_PO_WALL);
    base += sx * sy;
    for (uint32_t y = 0; y < sy; y++)
        for (uint32_t x = 0; x < sx; x++)
            W.put(base + y * sx + x, (spacing * x) + r + ox,
                  (spacing * y) + r + oy + bo * 2 * r + fy * 2.0f * r, (spacing * 0u) + r + oz,
                  CMAG_PO_WALL);
}

void init_peristalsis(const cmag_params& P, float* pos4, uint32_t n) {
    std::memset(pos4, 0, sizeof(float) * 4 * (size_t)n);
    Writer W{pos4, n};
    const float r = P.particle_radius, spacing = r * 2.0f;
    const int bo = P.boundary_offset;
    const uint32_t fx = P.fluid_size[0], fy = P.fluid_size[1], fz = P.fluid_size[2];
    const float halfX = P.grid_size[0] * r, halfZ = P.grid_size[2] * r;
    const float ox = P.world_origin[0], oz = P.world_origin[2];
    for (int z = 0; z < (int)fz; z++)
        for (int y = 0; y < (int)fy; y++)
            for (int x = 0; x < (int)fx; x++)
                W.put((z * fy * fx) + y * fx + x, (spacing * x) + r - halfX,
                      -1.0f * fy * r + (spacing * y) + r, (spacing * z) + r - halfZ, CMAG_PE_FLUID);
    uint32_t base = fx * fy * fz;
    const uint32_t sx = fx, sy = (uint32_t)bo;
    for (uint32_t y = 0; y < sy; y++)
        for (uint32_t x = 0; x < sx; x++) {
            float j = r * (2 * x + 1);
            W.put(base + y * sx + x, j + ox, -1.0f * fy * r - (spacing * y) - r, (spacing * 0u) + r + oz,
                  2.0f * (1.0f + y));
        }
    base += sx * sy;
    for (uint32_t y = 0; y < sy; y++)
        for (uint32_t x = 0; x < sx; x++) {
            float j = r * (2 * x + 1);
            W.put(base + y * sx + x, j + ox, fy * r + (spacing * y) + r, (spacing * 0u) + r + oz,
                  -2.0f * (1.0f + y));
        }
}

float peristalsis_lattice_sum(const cmag_params& P, const float* pos4, uint32_t n) {
    const uint32_t probe = P.grid_size[0] / 2;
    const float x = pos4[probe * 4 + 0], y = pos4[probe * 4 + 1];
    float sum = 0.0f;
    for (uint32_t i = 0; i < n; i++) {
        float rx = x - pos4[4 * (size_t)i], ry = y - pos4[4 * (size_t)i + 1];
        float dist = std::sqrt(rx * rx + ry * ry);
        float q = dist / P.smoothing_radius;
        float coeff = 7.0f / 4 / 3.141592654f / powf(P.smoothing_radius, 2);
        if (q < 2) sum += coeff * (powf(1 - 0.5f * q, 4) * (2 * q + 1));
    }
    return sum;
}

}  // namespace cmag
