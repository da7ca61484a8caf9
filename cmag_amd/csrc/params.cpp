// params.cpp - parameter blocks of the three reference systems and library
// utilities (errors, version).  Host-only code.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <limits>
#include <string>

#include "common.hpp"

namespace cmag {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
    set_error(msg);
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return CMAG_E_HIP;
}

int check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, what);
    return CMAG_OK;
}

static bool is_pow2(uint32_t v) { return v && !(v & (v - 1)); }

int validate_params(const cmag_params* p) {
    if (!p) return fail(CMAG_E_INVALID, "null parameter block");
    if (p->version != CMAG_PARAMS_VERSION) return fail(CMAG_E_INVALID, "parameter block version mismatch");
    for (int a = 0; a < 3; ++a) {
        if (!is_pow2(p->grid_size[a]))
            return fail(CMAG_E_INVALID, "grid_size must be a power of two on every axis");
        if ((1u << p->grid_bits[a]) != p->grid_size[a])
            return fail(CMAG_E_INVALID, "grid_bits inconsistent (call cmag_params_finalize)");
    }
    uint64_t cells = (uint64_t)p->grid_size[0] * p->grid_size[1] * p->grid_size[2];
    if (cells > (1ull << 31)) return fail(CMAG_E_INVALID, "more than 2^31 cells");
    if (p->cellcount < 0 || p->cellcount > 8) return fail(CMAG_E_INVALID, "cellcount outside [0,8]");
    if (p->system < CMAG_SYS_DAMBREAK || p->system > CMAG_SYS_PERISTALSIS)
        return fail(CMAG_E_INVALID, "unknown system");
    if (p->dim != 2 && p->dim != 3) return fail(CMAG_E_INVALID, "dim must be 2 or 3");
    if (p->system == CMAG_SYS_DAMBREAK && p->periodic_mask)
        return fail(CMAG_E_INVALID, "the DamBreak system has no periodic images");
    if (p->eos == CMAG_EOS_TAIT &&
        (p->gamma < 1.0f || p->gamma > 15.0f || p->gamma != std::floor(p->gamma)))
        return fail(CMAG_E_INVALID, "Tait gamma must be an integer in [1,15]");
    if (p->hash_order == CMAG_HASH_MORTON) {
        uint32_t all = p->morton_mask[0] | p->morton_mask[1] | p->morton_mask[2];
        if (all != (uint32_t)(cells - 1)) return fail(CMAG_E_INVALID, "Morton masks inconsistent");
    }
    return CMAG_OK;
}

static const float kPiF = 3.141592654f;  // CUDART_PI_F as the reference defines it

static void finalize_impl(cmag_params* p) {
    for (int a = 0; a < 3; ++a) {
        uint32_t b = 0;
        while ((1u << b) < p->grid_size[a] && b < 31) ++b;
        p->grid_bits[a] = b;
    }
    // Morton layout: the low morton_xbits bits of x first (bricks of 2^xbits
    // cells along x are contiguous), then bit b of every axis in x,y,z order
    // with x counted from bit xbits, axes without a bit b skipped
    p->morton_mask[0] = p->morton_mask[1] = p->morton_mask[2] = 0;
    uint32_t xb = p->morton_xbits < p->grid_bits[0] ? p->morton_xbits : p->grid_bits[0];
    p->morton_xbits = xb;
    uint32_t bits[3] = {p->grid_bits[0] - xb, p->grid_bits[1], p->grid_bits[2]};
    uint32_t bit = 0;
    for (uint32_t b = 0; b < xb; ++b) p->morton_mask[0] |= 1u << bit++;
    uint32_t maxb = bits[0];
    if (bits[1] > maxb) maxb = bits[1];
    if (bits[2] > maxb) maxb = bits[2];
    for (uint32_t b = 0; b < maxb; ++b)
        for (int a = 0; a < 3; ++a)
            if (b < bits[a]) p->morton_mask[a] |= 1u << bit++;
    // kernel normalisations with the reference's float operation order
    const float h = p->smoothing_radius;
    if (p->dim == 3) {
        p->kernel_coeff = 21.0f / (16.0f * kPiF * powf(h, 3.0f));
        p->grad_coeff = 21.0f / (8.0f * kPiF * powf(h, 4.0f));
    } else {
        p->kernel_coeff = 7.0f / 4 / kPiF / powf(h, 2.0f);               // fluid_kernel.cu:121
        if (p->system == CMAG_SYS_PERISTALSIS)
            p->grad_coeff = 7.0f / (2 * kPiF * powf(h, 3.0f));           // PeristalsisPressureForce.cu:38
        else
            p->grad_coeff = 7.0f / 2 / kPiF / powf(h, 3.0f);             // fluid_kernel.cu:222
    }
}

static void base_grid(cmag_params* p, uint32_t fx, uint32_t fy, uint32_t fz, int32_t bo,
                      uint32_t gx, uint32_t gy, uint32_t gz, float r) {
    std::memset(p, 0, sizeof(*p));
    p->version = CMAG_PARAMS_VERSION;
    p->hash_order = CMAG_HASH_LINEAR;
    p->dim = 2;
    p->cellcount = 3;
    p->fluid_size[0] = fx; p->fluid_size[1] = fy; p->fluid_size[2] = fz;
    p->grid_size[0] = gx; p->grid_size[1] = gy; p->grid_size[2] = gz;
    p->boundary_offset = bo;
    p->particle_radius = r;
    p->smoothing_radius = 3.0f * r;
    p->rest_density = 1000.0f;
    // worldOrigin = -(getHalfWorld*Size()) = -(gridSize * particleRadius)
    p->world_origin[0] = -(gx * r);
    p->world_origin[1] = -(gy * r);
    p->world_origin[2] = -(gz * r);
    float cs = r * 2.0f;
    p->cell_size[0] = p->cell_size[1] = p->cell_size[2] = cs;
    p->world_size[0] = gx * 2.0f * r;
    p->world_size[1] = gy * 2.0f * r;
    p->world_size[2] = gz * 2.0f * r;
    p->boundary_damping = -1.0f;
    p->alpha = 0.38f;
    p->beta_max = 1000.0f;
    const float inf = std::numeric_limits<float>::infinity();
    for (int a = 0; a < 3; ++a) { p->clamp_lo[a] = -inf; p->clamp_hi[a] = inf; }
    p->right_boundary = inf;
}

}  // namespace cmag

using namespace cmag;

extern "C" {

const char* cmag_last_error(void) { return g_last_error.c_str(); }

int cmag_version(void) { return 1; }

int cmag_params_size(void) { return (int)sizeof(cmag_params); }

int cmag_params_finalize(cmag_params* p) {
    if (!p) return fail(CMAG_E_INVALID, "null parameter block");
    p->version = CMAG_PARAMS_VERSION;
    finalize_impl(p);
    return validate_params(p);
}

uint32_t cmag_params_num_cells(const cmag_params* p) {
    return p ? p->grid_size[0] * p->grid_size[1] * p->grid_size[2] : 0u;
}

int cmag_params_key_bits(const cmag_params* p) {
    if (!p) return 0;
    return (int)(p->grid_bits[0] + p->grid_bits[1] + p->grid_bits[2]);
}

// DamBreakSystem::DamBreakSystem, DamBreak.Core/fluidSystem.cpp:17-77
int cmag_params_dambreak(cmag_params* p, uint32_t fx, uint32_t fy, uint32_t fz, int32_t bo,
                         uint32_t gx, uint32_t gy, uint32_t gz, float r) {
    if (!p) return fail(CMAG_E_INVALID, "null parameter block");
    base_grid(p, fx, fy, fz, bo, gx, gy, gz, r);
    p->system = CMAG_SYS_DAMBREAK;
    p->eos = CMAG_EOS_TAIT;
    p->gridpos_rule = CMAG_GRIDPOS_DIV_CELL;
    p->periodic_mask = 0;
    p->particle_mass = (float)(p->rest_density / 16327.6);  // float / double, as written
    p->right_boundary = p->world_origin[0] + (bo + fx) * 2 * r;
    p->gravity[0] = 0.0f; p->gravity[1] = -9.8f; p->gravity[2] = 0.0f;
    p->gamma = 7;
    p->B = 200 * p->rest_density * std::fabs(p->gravity[1]) * (2 * r * fy) / p->gamma;
    p->lj_D = 10 * p->gravity[1] * 2 * r * fy;
    p->lj_a = 1 * r;
    p->soundspeed = sqrtf(p->B * p->gamma / p->rest_density);
    p->delta_time = powf(10.0f, -4.0f);
    // integrate's borders, fluid_kernel.cu:352-353
    p->clamp_lo[0] = p->world_origin[0] + bo * 2 * r;
    p->clamp_lo[1] = p->world_origin[1] + bo * 2 * r;
    finalize_impl(p);
    return validate_params(p);
}

// PoiseuilleFlowSystem::PoiseuilleFlowSystem, Poiseuille.Core/poiseuilleFlowSystem.cpp:21-67
int cmag_params_poiseuille(cmag_params* p, uint32_t fx, uint32_t fy, uint32_t fz, int32_t bo,
                           uint32_t gx, uint32_t gy, uint32_t gz, float r) {
    if (!p) return fail(CMAG_E_INVALID, "null parameter block");
    base_grid(p, fx, fy, fz, bo, gx, gy, gz, r);
    p->system = CMAG_SYS_POISEUILLE;
    p->eos = CMAG_EOS_LINEAR_C2;
    p->gridpos_rule = CMAG_GRIDPOS_HALF_DIV_RADIUS;
    p->periodic_mask = 1;  // x images, poiseuilleFlowKernel.cu:157-163
    p->particle_mass = (float)(1000.0f / 3381320880.551724);
    p->gravity[0] = powf(10.0f, -4.0f); p->gravity[1] = 0.0f; p->gravity[2] = 0.0f;
    p->soundspeed = powf(10.0f, -4.0f);
    p->mu = powf(10.0f, -3.0f);
    p->delta_time = powf(10.0f, -4.0f);
    // getVelocityDiff's planes, poiseuilleFlowKernel.cu:186-187
    p->wall_bottom = p->world_origin[1] + bo * 2.0f * r;
    p->wall_top = p->wall_bottom + fy * 2.0f * r;
    finalize_impl(p);
    return validate_params(p);
}

// PeristalsisSystem::PeristalsisSystem, Peristalsis.Core/PeristalsisSystem.cpp:17-80
int cmag_params_peristalsis(cmag_params* p, float dt, uint32_t fx, uint32_t fy, uint32_t fz,
                            float amplitude, float wave_speed, float soundspeed, float gravx,
                            float gravy, float gravz, int32_t bo, uint32_t gx, uint32_t gy,
                            uint32_t gz, float r) {
    if (!p) return fail(CMAG_E_INVALID, "null parameter block");
    base_grid(p, fx, fy, fz, bo, gx, gy, gz, r);
    p->system = CMAG_SYS_PERISTALSIS;
    p->eos = CMAG_EOS_LINEAR_50C;
    p->gridpos_rule = CMAG_GRIDPOS_HALF_DIV_RADIUS;
    p->periodic_mask = 7;  // EvaluateShift on x, y and z
    p->amplitude = amplitude;
    p->wave_speed = wave_speed;
    p->particle_mass = 0.0f;  // set by Reset() from the lattice sum
    p->gravity[0] = gravx; p->gravity[1] = gravy; p->gravity[2] = gravz;
    p->soundspeed = soundspeed;
    p->mu = powf(10.0f, -3.0f);
    p->delta_time = dt;
    p->is_boundary_configuration = 1;
    p->wave_k = kPiF / (fx * r);  // GetWave, PeristalsisKernel.cuh:59-62
    finalize_impl(p);
    return validate_params(p);
}

}  // extern "C"
