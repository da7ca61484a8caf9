// common.hpp - shared host/device helpers of libcmag_hip.so (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "cmag_hip.h"

namespace cmag {

// ---- status / error reporting -------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define CMAG_HIP_TRY(call)                                      \
    do {                                                        \
        hipError_t cmag_e_ = (call);                            \
        if (cmag_e_ != hipSuccess) return ::cmag::hip_fail(cmag_e_, #call); \
    } while (0)

#define CMAG_TRY(call)                 \
    do {                               \
        int cmag_rc_ = (call);         \
        if (cmag_rc_ != CMAG_OK) return cmag_rc_; \
    } while (0)

inline hipStream_t as_stream(cmag_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// a launch left an error in the runtime (bad configuration, missing code object)
int check_launch(const char* what);

int validate_params(const cmag_params* p);

inline uint32_t div_up(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

// ---- device helpers -------------------------------------------------------

struct Cell {
    int x, y, z;
};

// calcGridPos (fluid_kernel.cu:19-25, poiseuilleFlowKernel.cu:18-24): the hash is
// bit-exact only with IEEE division and no contraction, which this function keeps.
__device__ __forceinline__ Cell grid_pos(const cmag_params& P, float x, float y, float z) {
#pragma clang fp contract(off)
    Cell c;
    if (P.gridpos_rule == CMAG_GRIDPOS_DIV_CELL) {
        c.x = (int)floorf((x - P.world_origin[0]) / P.cell_size[0]);
        c.y = (int)floorf((y - P.world_origin[1]) / P.cell_size[1]);
        c.z = (int)floorf((z - P.world_origin[2]) / P.cell_size[2]);
    } else {
        c.x = (int)floorf((x - P.world_origin[0]) * 0.5f / P.particle_radius);
        c.y = (int)floorf((y - P.world_origin[1]) * 0.5f / P.particle_radius);
        c.z = (int)floorf((z - P.world_origin[2]) * 0.5f / P.particle_radius);
    }
    return c;
}

// eighth (0..7) of a position's cell along x, from the quotient that
// grid_pos floors (the DIV_CELL rule); the sub-cell walk order buckets by it
constexpr int X_SUBS = 8;
__device__ __forceinline__ int x_sub(const cmag_params& P, float x) {
#pragma clang fp contract(off)
    const float u = (x - P.world_origin[0]) / P.cell_size[0];
    return min(X_SUBS - 1, (int)((u - floorf(u)) * (float)X_SUBS));
}

// software bit deposit of the low bits of v into the set bits of mask
__device__ __forceinline__ uint32_t deposit_bits(uint32_t v, uint32_t mask) {
    uint32_t out = 0;
    while (mask) {
        uint32_t low = mask & (0u - mask);
        if (v & 1u) out |= low;
        v >>= 1;
        mask &= mask - 1u;
    }
    return out;
}

// calcGridHash (fluid_kernel.cu:27-32): per-axis wrap with gridSize-1, then
// linear or Morton key
__device__ __forceinline__ uint32_t grid_hash(const cmag_params& P, int x, int y, int z) {
    uint32_t cx = (uint32_t)x & (P.grid_size[0] - 1u);
    uint32_t cy = (uint32_t)y & (P.grid_size[1] - 1u);
    uint32_t cz = (uint32_t)z & (P.grid_size[2] - 1u);
    if (P.hash_order == CMAG_HASH_LINEAR)
        return (cz * P.grid_size[1] + cy) * P.grid_size[0] + cx;
    return deposit_bits(cx, P.morton_mask[0]) | deposit_bits(cy, P.morton_mask[1]) |
           deposit_bits(cz, P.morton_mask[2]);
}

// Morton key arithmetic without re-interleaving: +1 / -1 on the bits of one
// axis (the masked carry wraps exactly like the gridSize-1 mask)
__device__ __forceinline__ uint32_t dilated_inc(uint32_t key, uint32_t mask) {
    return (((key | ~mask) + 1u) & mask) | (key & ~mask);
}
__device__ __forceinline__ uint32_t dilated_dec(uint32_t key, uint32_t mask) {
    return (((key & mask) - 1u) & mask) | (key & ~mask);
}

// image shift of a neighbour cell coordinate (PeristalsisUtil.cu:4-10)
__device__ __forceinline__ int image_shift(int c, uint32_t size) {
    return c < 0 ? -1 : (c > (int)size - 1 ? 1 : 0);
}

}  // namespace cmag
