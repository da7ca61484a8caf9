// integrate.hip - time integration and wall motion (gfx950).
//
//   integrate                   DamBreak.Core/fluid_kernel.cu:323-366 (clamp walls)
//   integratePoiseuilleSystemD  Poiseuille.Core/poiseuilleFlowKernel.cu:311-348 (x wrap)
//   computeCoordinatesD         Peristalsis.Core/PeristalsisIntegrate.cu:1-65
//                               (analytic sinusoidal walls, x wrap both ways)
//   configureBoundaryD          Peristalsis.Core/PeristalsisKernel.cu:144-171
//   shift/removeRightBoundaryD  DamBreak.Core/fluid_kernel.cu:294-321
// Streaming kernels: one float4 load/store per array per particle.
#include "common.hpp"
#include "kernels.hpp"

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

template <int SYS>
__global__ __launch_bounds__(256) void integrate_kernel(const cmag_params P, float4* __restrict__ pos,
                                                         float4* __restrict__ vel,
                                                         float4* __restrict__ velLF,
                                                         const float4* __restrict__ acc,
                                                         const float4* __restrict__ presF,
                                                         float wave_time, uint32_t n) {
#pragma clang fp contract(off)
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = pos[i];
    if (SYS == CMAG_SYS_DAMBREAK && p.w != CMAG_DB_FLUID) return;
    if (SYS == CMAG_SYS_POISEUILLE && p.w == CMAG_PO_WALL) return;
    if (SYS == CMAG_SYS_PERISTALSIS && p.w != 0.0f) {
        if (P.is_boundary_configuration) return;
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

int launch_integrate(const cmag_params& P, float* pos, float* vel, float* velLF, const float* acc,
                     const float* presF, float wave_time, uint32_t n, hipStream_t s) {
    if (n == 0) return CMAG_OK;
    dim3 g(div_up(n, 256)), b(256);
    auto f4 = [](float* q) { return reinterpret_cast<float4*>(q); };
    auto c4 = [](const float* q) { return reinterpret_cast<const float4*>(q); };
    if (P.system == CMAG_SYS_DAMBREAK)
        hipLaunchKernelGGL(integrate_kernel<CMAG_SYS_DAMBREAK>, g, b, 0, s, P, f4(pos), f4(vel), f4(velLF), c4(acc), c4(presF), wave_time, n);
    else if (P.system == CMAG_SYS_POISEUILLE)
        hipLaunchKernelGGL(integrate_kernel<CMAG_SYS_POISEUILLE>, g, b, 0, s, P, f4(pos), f4(vel), f4(velLF), c4(acc), c4(presF), wave_time, n);
    else
        hipLaunchKernelGGL(integrate_kernel<CMAG_SYS_PERISTALSIS>, g, b, 0, s, P, f4(pos), f4(vel), f4(velLF), c4(acc), c4(presF), wave_time, n);
    return check_launch("integrate_kernel");
}

__global__ __launch_bounds__(256) void configure_boundary_kernel(const cmag_params P, float4* __restrict__ pos,
                                                                  float wave_height, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = pos[i];
    if (p.w > 0.0f || p.w < 0.0f) pos[i] = make_float4(p.x, wall_y(P, p.x, p.w, wave_height, 0.0f), p.z, p.w);
}

int launch_configure_boundary(const cmag_params& P, float* pos, float wave_height, uint32_t n,
                              hipStream_t s) {
    if (n == 0) return CMAG_OK;
    hipLaunchKernelGGL(configure_boundary_kernel, dim3(div_up(n, 256)), dim3(256), 0, s, P,
                       reinterpret_cast<float4*>(pos), wave_height, n);
    return check_launch("configure_boundary_kernel");
}

template <bool REMOVE>
__global__ __launch_bounds__(256) void right_boundary_kernel(const cmag_params P, float4* __restrict__ pos,
                                                              uint32_t n) {
#pragma clang fp contract(off)
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = pos[i];
    if (p.w != CMAG_DB_RIGHT_FIRST && p.w != CMAG_DB_RIGHT_SECOND) return;
    const float width = (float)(P.fluid_size[0] * 2u) * P.particle_radius;
    if (REMOVE) {
        const float height = (float)(P.grid_size[1] * 2u) * P.particle_radius;
        pos[i] = make_float4(width, height, p.z, p.w);
    } else {
        pos[i] = make_float4(p.x + width, p.y, p.z, p.w);
    }
}

int launch_shift_right_boundary(const cmag_params& P, float* pos, uint32_t n, hipStream_t s) {
    if (n == 0) return CMAG_OK;
    hipLaunchKernelGGL(right_boundary_kernel<false>, dim3(div_up(n, 256)), dim3(256), 0, s, P,
                       reinterpret_cast<float4*>(pos), n);
    return check_launch("right_boundary_kernel<shift>");
}

int launch_remove_right_boundary(const cmag_params& P, float* pos, uint32_t n, hipStream_t s) {
    if (n == 0) return CMAG_OK;
    hipLaunchKernelGGL(right_boundary_kernel<true>, dim3(div_up(n, 256)), dim3(256), 0, s, P,
                       reinterpret_cast<float4*>(pos), n);
    return check_launch("right_boundary_kernel<remove>");
}

}  // namespace cmag

using namespace cmag;

extern "C" {

int cmag_integrate(const cmag_params* p, float* pos, float* vel, float* velLF, const float* acc,
                   const float* presF, float wave_time, uint32_t n, cmag_stream_t stream) {
    CMAG_TRY(validate_params(p));
    if (n && (!pos || !vel || !velLF || !acc))
        return fail(CMAG_E_INVALID, "cmag_integrate: null pointer");
    if (n && p->system == CMAG_SYS_PERISTALSIS && !presF)
        return fail(CMAG_E_INVALID, "cmag_integrate: Peristalsis needs the pressure force");
    return launch_integrate(*p, pos, vel, velLF, acc, presF, wave_time, n, as_stream(stream));
}

int cmag_configure_boundary(const cmag_params* p, float* pos, float wave_height, uint32_t n,
                            cmag_stream_t stream) {
    CMAG_TRY(validate_params(p));
    if (n && !pos) return fail(CMAG_E_INVALID, "cmag_configure_boundary: null pointer");
    return launch_configure_boundary(*p, pos, wave_height, n, as_stream(stream));
}

int cmag_shift_right_boundary(const cmag_params* p, float* pos, uint32_t n, cmag_stream_t stream) {
    CMAG_TRY(validate_params(p));
    if (n && !pos) return fail(CMAG_E_INVALID, "cmag_shift_right_boundary: null pointer");
    return launch_shift_right_boundary(*p, pos, n, as_stream(stream));
}

int cmag_remove_right_boundary(const cmag_params* p, float* pos, uint32_t n, cmag_stream_t stream) {
    CMAG_TRY(validate_params(p));
    if (n && !pos) return fail(CMAG_E_INVALID, "cmag_remove_right_boundary: null pointer");
    return launch_remove_right_boundary(*p, pos, n, as_stream(stream));
}

}  // extern "C"
