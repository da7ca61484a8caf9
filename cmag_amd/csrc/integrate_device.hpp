// integrate_device.hpp - per-particle time integration and wall motion,
// shared by the single-system kernels (integrate.hip) and the batched
// persistent kernels (batch.hip).
//
//   integrate                   DamBreak.Core/fluid_kernel.cu:323-366 (clamp walls)
//   integratePoiseuilleSystemD  Poiseuille.Core/poiseuilleFlowKernel.cu:311-348 (x wrap)
//   computeCoordinatesD         Peristalsis.Core/PeristalsisIntegrate.cu:1-65
//                               (analytic sinusoidal walls, x wrap both ways)
#pragma once

#include "common.hpp"

namespace cmag {

// GetWave, PeristalsisKernel.cuh:59-62
__device__ __forceinline__ float wave(const cmag_params& P, float x, float t) {
    return sinf(P.wave_k * ((x - P.world_origin[0]) - P.wave_speed * t));
}

__device__ __forceinline__ float wall_y(const cmag_params& P, float x, float w, float A, float t) {
#pragma clang fp contract(off)
    const float ny = (float)P.fluid_size[1], r = P.particle_radius;
    if (w > 0.0f) return -1.0f * ny * r + A * wave(P, x, t) - r * (w - 1.0f);
    return ny * r - A * wave(P, x, t) + r * (-w - 1.0f);
}

// time derivative of wall_y (only with wall_velocity set; not in the reference)
__device__ __forceinline__ float wall_vy(const cmag_params& P, float x, float w, float A, float t) {
    const float arg = P.wave_k * ((x - P.world_origin[0]) - P.wave_speed * t);
    const float d = A * cosf(arg) * P.wave_k * (-1.0f * P.wave_speed);
    return w > 0.0f ? d : -d;
}

// one particle of integrate; is_config is the Peristalsis boundary-configuration flag
template <int SYS>
__device__ __forceinline__ void integrate_body(const cmag_params& P, uint32_t i, float4* __restrict__ pos,
                                               float4* __restrict__ vel, float4* __restrict__ velLF,
                                               const float4* __restrict__ acc,
                                               const float4* __restrict__ presF, float wave_time,
                                               int is_config) {
#pragma clang fp contract(off)
    const float4 p = pos[i];
    if (SYS == CMAG_SYS_DAMBREAK && p.w != CMAG_DB_FLUID) return;
    if (SYS == CMAG_SYS_POISEUILLE && p.w == CMAG_PO_WALL) return;
    if (SYS == CMAG_SYS_PERISTALSIS && p.w != 0.0f) {
        if (is_config) return;
        pos[i] = make_float4(p.x, wall_y(P, p.x, p.w, P.amplitude, wave_time), p.z, p.w);
        if (P.wall_velocity) {
            const float vy = wall_vy(P, p.x, p.w, P.amplitude, wave_time);
            vel[i] = make_float4(0.0f, vy, 0.0f, vel[i].w);
            velLF[i] = make_float4(0.0f, vy, 0.0f, velLF[i].w);
        }
        return;
    }
    const float dt = P.delta_time;
    const float4 v = vel[i];
    const float4 a = acc[i];
    const float4 l = velLF[i];
    float ax = P.gravity[0] + a.x, ay = P.gravity[1] + a.y, az = P.gravity[2] + a.z;
    if (SYS == CMAG_SYS_PERISTALSIS) {
        const float4 pr = presF[i];
        ax += pr.x; ay += pr.y; az += pr.z;
    }
    float nx = v.x + ax * dt, ny = v.y + ay * dt, nz = v.z + az * dt;
    const float lx = (v.x + nx) * 0.5f, ly = (v.y + ny) * 0.5f, lz = (v.z + nz) * 0.5f;
    float px = p.x + nx * dt, py = p.y + ny * dt, pz = p.z + nz * dt;
    if (SYS == CMAG_SYS_DAMBREAK) {
        const float damp = P.boundary_damping;
        if (px < P.clamp_lo[0]) { px = P.clamp_lo[0]; nx *= damp; }
        if (px > P.right_boundary) { px = P.right_boundary; nx *= damp; }
        if (py < P.clamp_lo[1]) { py = P.clamp_lo[1]; ny *= damp; }
        if (py > P.clamp_hi[1]) { py = P.clamp_hi[1]; ny *= damp; }
        if (pz < P.clamp_lo[2]) { pz = P.clamp_lo[2]; nz *= damp; }
        if (pz > P.clamp_hi[2]) { pz = P.clamp_hi[2]; nz *= damp; }
    } else {
        const float lo = P.world_origin[0];
        const float hi = lo + P.world_size[0];
        if (px > hi) px -= P.world_size[0];
        if (SYS == CMAG_SYS_PERISTALSIS && px < lo) px += P.world_size[0];
    }
    pos[i] = make_float4(px, py, pz, p.w);
    vel[i] = make_float4(nx, ny, nz, v.w);
    velLF[i] = make_float4(lx, ly, lz, l.w);
}

}  // namespace cmag
