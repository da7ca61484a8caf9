// integrate.hip - time integration and wall motion (gfx950).  Per-particle
// bodies (integrate, wall sinusoid) live in integrate_device.hpp.
//
//   integrate                   DamBreak.Core/fluid_kernel.cu:323-366 (clamp walls)
//   integratePoiseuilleSystemD  Poiseuille.Core/poiseuilleFlowKernel.cu:311-348 (x wrap)
//   computeCoordinatesD         Peristalsis.Core/PeristalsisIntegrate.cu:1-65
//                               (analytic sinusoidal walls, x wrap both ways)
//   configureBoundaryD          Peristalsis.Core/PeristalsisKernel.cu:144-171
//   shift/removeRightBoundaryD  DamBreak.Core/fluid_kernel.cu:294-321
// Streaming kernels: one float4 load/store per array per particle.
#include "common.hpp"
#include "integrate_device.hpp"
#include "kernels.hpp"
#include "step_state.hpp"

namespace cmag {

template <int SYS>
__global__ __launch_bounds__(256) void integrate_kernel(const cmag_params P, float4* __restrict__ pos,
                                                         float4* __restrict__ vel,
                                                         float4* __restrict__ velLF,
                                                         const float4* __restrict__ acc,
                                                         const float4* __restrict__ presF,
                                                         float wave_time, const StepState* __restrict__ st,
                                                         uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (st) integrate_body<SYS>(P, i, pos, vel, velLF, acc, presF, st->t_w, st->is_config);
    else integrate_body<SYS>(P, i, pos, vel, velLF, acc, presF, wave_time, P.is_boundary_configuration);
}

int launch_integrate(const cmag_params& P, float* pos, float* vel, float* velLF, const float* acc,
                     const float* presF, float wave_time, uint32_t n, hipStream_t s, const StepState* st) {
    if (n == 0) return CMAG_OK;
    dim3 g(div_up(n, 256)), b(256);
    auto f4 = [](float* q) { return reinterpret_cast<float4*>(q); };
    auto c4 = [](const float* q) { return reinterpret_cast<const float4*>(q); };
    if (P.system == CMAG_SYS_DAMBREAK)
        hipLaunchKernelGGL(integrate_kernel<CMAG_SYS_DAMBREAK>, g, b, 0, s, P, f4(pos), f4(vel), f4(velLF), c4(acc), c4(presF), wave_time, st, n);
    else if (P.system == CMAG_SYS_POISEUILLE)
        hipLaunchKernelGGL(integrate_kernel<CMAG_SYS_POISEUILLE>, g, b, 0, s, P, f4(pos), f4(vel), f4(velLF), c4(acc), c4(presF), wave_time, st, n);
    else
        hipLaunchKernelGGL(integrate_kernel<CMAG_SYS_PERISTALSIS>, g, b, 0, s, P, f4(pos), f4(vel), f4(velLF), c4(acc), c4(presF), wave_time, st, n);
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

__global__ __launch_bounds__(64) void step_prologue_kernel(const cmag_params P, StepState* __restrict__ state) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    StepState st = *state;
    advance_step_state(P, st);
    *state = st;
}

int launch_step_prologue(const cmag_params& P, StepState* state, hipStream_t s) {
    hipLaunchKernelGGL(step_prologue_kernel, dim3(1), dim3(64), 0, s, P, state);
    return check_launch("step_prologue_kernel");
}

// configureBoundaryD (PeristalsisKernel.cu:144-171) when this step's state
// record asks for it; a no-op step otherwise
__global__ __launch_bounds__(256) void configure_boundary_state_kernel(const cmag_params P, float4* __restrict__ pos,
                                                                        const StepState* __restrict__ st,
                                                                        uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !st->configure_now) return;
    const float4 p = pos[i];
    if (p.w > 0.0f || p.w < 0.0f)
        pos[i] = make_float4(p.x, wall_y(P, p.x, p.w, st->configure_height, 0.0f), p.z, p.w);
}

int launch_configure_boundary_from_state(const cmag_params& P, float* pos, const StepState* state, uint32_t n,
                                         hipStream_t s) {
    if (n == 0) return CMAG_OK;
    hipLaunchKernelGGL(configure_boundary_state_kernel, dim3(div_up(n, 256)), dim3(256), 0, s, P,
                       reinterpret_cast<float4*>(pos), state, n);
    return check_launch("configure_boundary_state_kernel");
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
