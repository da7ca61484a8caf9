// sph.hip - neighbour passes of the WCSPH step (gfx950, wave64, FP32).
//
// One thread per particle in sorted order; the stencil of (2*cellcount+1)^dim
// cells is walked in the reference's z, y, x order and each cell's particles in
// sorted order, so the sums accumulate in the reference's order.  Candidates
// are culled on r^2 < 4h^2 before any square root; inner-loop reciprocals use
// the hardware v_rcp/v_sqrt (1 ulp), the per-particle divisions stay IEEE.
//
//   density       fluid_kernel.cu:99-169, poiseuilleFlowKernel.cu:99-182,
//                 PeristalsisDensity.cu:3-79
//   force         fluid_kernel.cu:172-292 (DamBreak: pressure + Monaghan AV,
//                 Lennard-Jones walls), poiseuilleFlowKernel.cu:184-309
//                 (pressure + Morris viscosity with wall extrapolation)
//   pressure /    PeristalsisPressureForce.cu:3-90,
//   viscous       PeristalsisViscousForce.cu:143-241
#include "common.hpp"
#include "kernels.hpp"

namespace cmag {

__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// integer power by squaring (gamma is an integer in [1,15])
__device__ __forceinline__ float pow_int(float x, int e) {
    float r = 1.0f, b = x;
    while (e) {
        if (e & 1) r *= b;
        b *= b;
        e >>= 1;
    }
    return r;
}

__device__ __forceinline__ float eos(const cmag_params& P, float dens) {
    if (P.eos == CMAG_EOS_TAIT) return P.B * (pow_int(dens / P.rest_density, (int)P.gamma) - 1.0f);
    if (P.eos == CMAG_EOS_LINEAR_C2) return P.soundspeed * P.soundspeed * dens;
    return 50.0f * P.soundspeed * dens;
}

// walk the neighbour cells of gp; f(start, end, swx, swy, swz) with the image
// offsets s*worldSize of each axis (zero on non-periodic axes)
template <int DIM, typename F>
__device__ __forceinline__ void for_each_cell(const cmag_params& P, const Cell& gp,
                                              const uint32_t* __restrict__ cellStart,
                                              const uint32_t* __restrict__ cellEnd, uint32_t n, F&& f) {
    const int cc = P.cellcount;
    const int zlo = DIM == 2 ? 0 : -cc, zhi = DIM == 2 ? 0 : cc;
    for (int dz = zlo; dz <= zhi; ++dz) {
        const int nz = gp.z + dz;
        const float swz = (P.periodic_mask & 4) ? (float)image_shift(nz, P.grid_size[2]) * P.world_size[2] : 0.0f;
        for (int dy = -cc; dy <= cc; ++dy) {
            const int ny = gp.y + dy;
            const float swy = (P.periodic_mask & 2) ? (float)image_shift(ny, P.grid_size[1]) * P.world_size[1] : 0.0f;
            for (int dx = -cc; dx <= cc; ++dx) {
                const int nx = gp.x + dx;
                const float swx = (P.periodic_mask & 1) ? (float)image_shift(nx, P.grid_size[0]) * P.world_size[0] : 0.0f;
                const uint32_t h = grid_hash(P, nx, ny, nz);
                const uint32_t start = cellStart[h];
                if (start == CMAG_CELL_EMPTY) continue;
                // ranges are clamped to the particle count: consistent cell
                // arrays never need it, inconsistent ones cannot read past n
                f(start, min(cellEnd[h], n), swx, swy, swz);
            }
        }
    }
}

// ---------------------------------------------------------------------------
template <int SYS, int DIM>
__global__ __launch_bounds__(256) void density_kernel(const cmag_params P, float4* __restrict__ measures,
                                                       const float4* __restrict__ sortedPos,
                                                       const uint32_t* __restrict__ cellStart,
                                                       const uint32_t* __restrict__ cellEnd, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 pi = sortedPos[i];
    float2* m2 = reinterpret_cast<float2*>(measures);
    if (SYS == CMAG_SYS_POISEUILLE && pi.w == CMAG_PO_WALL) {
        m2[2 * (size_t)i] = make_float2(P.rest_density, eos(P, P.rest_density));
        return;
    }
    const Cell gp = grid_pos(P, pi.x, pi.y, pi.z);
    const float h = P.smoothing_radius;
    const float inv_h = 1.0f / h;
    const float r2max = 4.0f * h * h;
    float sum = 0.0f;
    for_each_cell<DIM>(P, gp, cellStart, cellEnd, n,
                       [&](uint32_t start, uint32_t end, float swx, float swy, float swz) {
                           float part = 0.0f;
                           for (uint32_t j = start; j < end; ++j) {
                               const float4 pj = sortedPos[j];
                               const float rx = pi.x - (pj.x + swx);
                               const float ry = pi.y - (pj.y + swy);
                               const float rz = pi.z - (pj.z + swz);
                               const float r2 = rx * rx + ry * ry + rz * rz;
                               if (r2 < r2max) {
                                   const float q = fsqrt(r2) * inv_h;
                                   const float t = 1.0f - 0.5f * q;
                                   const float t2 = t * t;
                                   part += t2 * t2 * (2.0f * q + 1.0f);
                               }
                           }
                           sum += part;
                       });
    const float dens = sum * P.kernel_coeff * P.particle_mass;
    if (SYS == CMAG_SYS_PERISTALSIS)
        measures[i] = make_float4(dens, eos(P, dens), 0.0f, pi.w);
    else
        m2[2 * (size_t)i] = make_float2(dens, eos(P, dens));
}

// getVelocityDiff, poiseuilleFlowKernel.cu:184-208
__device__ __forceinline__ float3 vel_diff_poiseuille(const cmag_params& P, const float4& vi,
                                                      const float4& pi, const float4& vj,
                                                      const float4& pj) {
    if (pj.w == CMAG_PO_WALL && pj.y > P.wall_top) {
        float beta = fminf(P.beta_max, 1.0f + (pj.y - P.wall_top) / (P.wall_top - pi.y));
        return make_float3(beta * vi.x, beta * vi.y, beta * vi.z);
    }
    if (pj.w == CMAG_PO_WALL && pj.y < P.wall_bottom) {
        float beta = fminf(P.beta_max, 1.0f + (P.wall_bottom - pj.y) / (pi.y - P.wall_bottom));
        return make_float3(beta * vi.x, beta * vi.y, beta * vi.z);
    }
    return make_float3(vi.x - vj.x, vi.y - vj.y, vi.z - vj.z);
}

template <int SYS, int DIM>
__global__ __launch_bounds__(256) void force_kernel(const cmag_params P, float4* __restrict__ acc,
                                                     const float4* __restrict__ measures,
                                                     const float4* __restrict__ sortedPos,
                                                     const float4* __restrict__ sortedVel,
                                                     const uint32_t* __restrict__ index,
                                                     const uint32_t* __restrict__ cellStart,
                                                     const uint32_t* __restrict__ cellEnd, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 pi = sortedPos[i];
    if (SYS == CMAG_SYS_DAMBREAK && pi.w != CMAG_DB_FLUID) return;
    const float4 vi = sortedVel[i];
    const float4 mi = measures[i];
    const float dens = mi.x;
    const float pterm_i = mi.y / (dens * dens);
    const Cell gp = grid_pos(P, pi.x, pi.y, pi.z);
    const float h = P.smoothing_radius;
    const float inv_h = 1.0f / h;
    const float r2max = 4.0f * h * h;
    const float m = P.particle_mass;
    const float gc = P.grad_coeff;
    const float av_k = 2.0f * P.alpha * h * P.soundspeed;  // nu * (rho_i + rho_j)
    const float av_eps = 0.001f * h * h;
    const float visc_k = m * (P.mu + P.mu);
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    for_each_cell<DIM>(P, gp, cellStart, cellEnd, n, [&](uint32_t start, uint32_t end, float swx,
                                                      float swy, float swz) {
        float tx = 0.0f, ty = 0.0f, tz = 0.0f;
        for (uint32_t j = start; j < end; ++j) {
            if (j == i) continue;
            const float4 pj = sortedPos[j];
            const float rx = pi.x - (pj.x + swx);
            const float ry = pi.y - (pj.y + swy);
            const float rz = pi.z - (pj.z + swz);
            const float r2 = rx * rx + ry * ry + rz * rz;
            if (SYS == CMAG_SYS_DAMBREAK && pj.w != CMAG_DB_FLUID) {
                // Lennard-Jones wall term, every wall candidate (no cutoff)
                const float dist = fsqrt(r2);
                const float ad = P.lj_a * frcp(dist);
                const float ad2 = ad * ad;
                const float ad6 = ad2 * ad2 * ad2;
                const float sc = P.lj_D * (ad6 * ad6 - ad6) * frcp(dist * dist);
                tx += sc * rx; ty += sc * ry; tz += sc * rz;
                continue;
            }
            if (!(r2 < r2max)) continue;
            const float4 mj = measures[j];
            const float4 vj = sortedVel[j];
            const float dist = fsqrt(r2);
            const float inv_d = frcp(dist);
            const float q = dist * inv_h;
            const float t = 1.0f - 0.5f * q;
            const float t2 = t * t;
            const float temp = gc * (t2 * t2 - t2 * t * (2.0f * q + 1.0f));
            const float inv_dj = frcp(mj.x);
            float s = pterm_i + mj.y * inv_dj * inv_dj;
            if (SYS == CMAG_SYS_DAMBREAK) {
                const float vr = (vi.x - vj.x) * rx + (vi.y - vj.y) * ry + (vi.z - vj.z) * rz;
                if (vr < 0.0f) s += -av_k * frcp(dens + mj.x) * vr * frcp(r2 + av_eps);
                const float k = -m * s * inv_d * temp;
                tx += k * rx; ty += k * ry; tz += k * rz;
            } else {
                const float3 V = vel_diff_poiseuille(P, vi, pi, vj, pj);
                const float k = -m * s * inv_d * temp;
                const float kv = visc_k * frcp(dens * mj.x) * inv_d * temp;
                tx += k * rx + kv * V.x;
                ty += k * ry + kv * V.y;
                tz += k * rz + kv * V.z;
            }
        }
        fx += tx; fy += ty; fz += tz;
    });
    const uint32_t k = index[i];
    if (k < n) acc[k] = make_float4(fx, fy, fz, 0.0f);
}

// Peristalsis: pressure and viscous passes; the fused variant produces both
// with one neighbour walk (same per-pair arithmetic as the split kernels)
template <int DIM, bool WANT_VISC, bool WANT_PRES>
__global__ __launch_bounds__(256) void peristalsis_force_kernel(
    const cmag_params P, float4* __restrict__ viscF, float4* __restrict__ presF,
    const float4* __restrict__ measures, const float4* __restrict__ sortedPos,
    const float4* __restrict__ sortedVel, const uint32_t* __restrict__ index,
    const uint32_t* __restrict__ cellStart, const uint32_t* __restrict__ cellEnd, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 pi = sortedPos[i];
    const float4 mi = measures[i];
    const float dens = mi.x;
    const float pterm_i = mi.y / (dens * dens);
    float4 vi = make_float4(0.f, 0.f, 0.f, 0.f);
    if (WANT_VISC) vi = sortedVel[i];
    const Cell gp = grid_pos(P, pi.x, pi.y, pi.z);
    const float h = P.smoothing_radius;
    const float inv_h = 1.0f / h;
    const float r2max = 4.0f * h * h;
    const float m = P.particle_mass;
    const float gc = P.grad_coeff;
    const float visc_k = m * (P.mu + P.mu);
    float vx = 0.f, vy = 0.f, vz = 0.f, px = 0.f, py = 0.f, pz = 0.f;
    for_each_cell<DIM>(P, gp, cellStart, cellEnd, n, [&](uint32_t start, uint32_t end, float swx,
                                                      float swy, float swz) {
        float tvx = 0.f, tvy = 0.f, tvz = 0.f, tpx = 0.f, tpy = 0.f, tpz = 0.f;
        for (uint32_t j = start; j < end; ++j) {
            if (j == i) continue;
            const float4 pj = sortedPos[j];
            const float rx = pi.x - (pj.x + swx);
            const float ry = pi.y - (pj.y + swy);
            const float rz = pi.z - (pj.z + swz);
            const float r2 = rx * rx + ry * ry + rz * rz;
            if (!(r2 < r2max)) continue;
            const float4 mj = measures[j];
            const float dist = fsqrt(r2);
            const float inv_d = frcp(dist);
            const float q = dist * inv_h;
            const float t = 1.0f - 0.5f * q;
            const float t2 = t * t;
            const float temp = gc * (t2 * t2 - t2 * t * (2.0f * q + 1.0f));
            if (WANT_PRES) {
                const float inv_dj = frcp(mj.x);
                const float k = -m * temp * (pterm_i + mj.y * inv_dj * inv_dj) * inv_d;
                tpx += k * rx; tpy += k * ry; tpz += k * rz;
            }
            if (WANT_VISC) {
                const float4 vj = sortedVel[j];
                const float kv = visc_k * temp * frcp(dens * mj.x * dist);
                tvx += kv * (vi.x - vj.x); tvy += kv * (vi.y - vj.y); tvz += kv * (vi.z - vj.z);
            }
        }
        vx += tvx; vy += tvy; vz += tvz;
        px += tpx; py += tpy; pz += tpz;
    });
    const uint32_t k = index[i];
    if (k >= n) return;
    if (WANT_VISC) viscF[k] = make_float4(vx, vy, vz, 0.0f);
    if (WANT_PRES) presF[k] = make_float4(px, py, pz, 0.0f);
}

// ---- launchers --------------------------------------------------------------
#define CMAG_DISPATCH_SYS_DIM(P, KERNEL, ...)                                        \
    do {                                                                            \
        if ((P).dim == 3) {                                                         \
            if ((P).system == CMAG_SYS_DAMBREAK) KERNEL<CMAG_SYS_DAMBREAK, 3>(__VA_ARGS__);         \
            else if ((P).system == CMAG_SYS_POISEUILLE) KERNEL<CMAG_SYS_POISEUILLE, 3>(__VA_ARGS__); \
            else KERNEL<CMAG_SYS_PERISTALSIS, 3>(__VA_ARGS__);                      \
        } else {                                                                    \
            if ((P).system == CMAG_SYS_DAMBREAK) KERNEL<CMAG_SYS_DAMBREAK, 2>(__VA_ARGS__);         \
            else if ((P).system == CMAG_SYS_POISEUILLE) KERNEL<CMAG_SYS_POISEUILLE, 2>(__VA_ARGS__); \
            else KERNEL<CMAG_SYS_PERISTALSIS, 2>(__VA_ARGS__);                      \
        }                                                                           \
    } while (0)

template <int SYS, int DIM>
static void go_density(const cmag_params& P, float* measures, const float* sortedPos,
                       const uint32_t* cellStart, const uint32_t* cellEnd, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL((density_kernel<SYS, DIM>), dim3(div_up(n, 256)), dim3(256), 0, s, P,
                       reinterpret_cast<float4*>(measures), reinterpret_cast<const float4*>(sortedPos),
                       cellStart, cellEnd, n);
}

int launch_density(const cmag_params& P, float* measures, const float* sortedPos,
                   const uint32_t* cellStart, const uint32_t* cellEnd, uint32_t n, hipStream_t s) {
    if (n == 0) return CMAG_OK;
    CMAG_DISPATCH_SYS_DIM(P, go_density, P, measures, sortedPos, cellStart, cellEnd, n, s);
    return check_launch("density_kernel");
}

template <int SYS, int DIM>
static void go_force(const cmag_params& P, float* acc, const float* measures, const float* sortedPos,
                     const float* sortedVel, const uint32_t* index, const uint32_t* cellStart,
                     const uint32_t* cellEnd, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL((force_kernel<SYS, DIM>), dim3(div_up(n, 256)), dim3(256), 0, s, P,
                       reinterpret_cast<float4*>(acc), reinterpret_cast<const float4*>(measures),
                       reinterpret_cast<const float4*>(sortedPos),
                       reinterpret_cast<const float4*>(sortedVel), index, cellStart, cellEnd, n);
}

int launch_force(const cmag_params& P, float* acc, const float* measures, const float* sortedPos,
                 const float* sortedVel, const uint32_t* index, const uint32_t* cellStart,
                 const uint32_t* cellEnd, uint32_t n, hipStream_t s) {
    if (n == 0) return CMAG_OK;
    if (P.system == CMAG_SYS_PERISTALSIS)
        return fail(CMAG_E_INVALID, "Peristalsis uses the split pressure/viscous passes");
    if (P.dim == 3) {
        if (P.system == CMAG_SYS_DAMBREAK) go_force<CMAG_SYS_DAMBREAK, 3>(P, acc, measures, sortedPos, sortedVel, index, cellStart, cellEnd, n, s);
        else go_force<CMAG_SYS_POISEUILLE, 3>(P, acc, measures, sortedPos, sortedVel, index, cellStart, cellEnd, n, s);
    } else {
        if (P.system == CMAG_SYS_DAMBREAK) go_force<CMAG_SYS_DAMBREAK, 2>(P, acc, measures, sortedPos, sortedVel, index, cellStart, cellEnd, n, s);
        else go_force<CMAG_SYS_POISEUILLE, 2>(P, acc, measures, sortedPos, sortedVel, index, cellStart, cellEnd, n, s);
    }
    return check_launch("force_kernel");
}

template <bool V, bool PR>
static void go_peri(const cmag_params& P, float* viscF, float* presF, const float* measures,
                    const float* sortedPos, const float* sortedVel, const uint32_t* index,
                    const uint32_t* cellStart, const uint32_t* cellEnd, uint32_t n, hipStream_t s) {
    if (P.dim == 3)
        hipLaunchKernelGGL((peristalsis_force_kernel<3, V, PR>), dim3(div_up(n, 256)), dim3(256), 0, s, P,
                           reinterpret_cast<float4*>(viscF), reinterpret_cast<float4*>(presF),
                           reinterpret_cast<const float4*>(measures), reinterpret_cast<const float4*>(sortedPos),
                           reinterpret_cast<const float4*>(sortedVel), index, cellStart, cellEnd, n);
    else
        hipLaunchKernelGGL((peristalsis_force_kernel<2, V, PR>), dim3(div_up(n, 256)), dim3(256), 0, s, P,
                           reinterpret_cast<float4*>(viscF), reinterpret_cast<float4*>(presF),
                           reinterpret_cast<const float4*>(measures), reinterpret_cast<const float4*>(sortedPos),
                           reinterpret_cast<const float4*>(sortedVel), index, cellStart, cellEnd, n);
}

int launch_pressure_force(const cmag_params& P, float* presF, const float* measures,
                          const float* sortedPos, const uint32_t* index, const uint32_t* cellStart,
                          const uint32_t* cellEnd, uint32_t n, hipStream_t s) {
    if (n == 0) return CMAG_OK;
    go_peri<false, true>(P, nullptr, presF, measures, sortedPos, nullptr, index, cellStart, cellEnd, n, s);
    return check_launch("peristalsis_force_kernel<pressure>");
}

int launch_viscous_force(const cmag_params& P, float* viscF, const float* measures,
                         const float* sortedPos, const float* sortedVel, const uint32_t* index,
                         const uint32_t* cellStart, const uint32_t* cellEnd, uint32_t n,
                         hipStream_t s) {
    if (n == 0) return CMAG_OK;
    go_peri<true, false>(P, viscF, nullptr, measures, sortedPos, sortedVel, index, cellStart, cellEnd, n, s);
    return check_launch("peristalsis_force_kernel<viscous>");
}

int launch_peristalsis_forces(const cmag_params& P, float* viscF, float* presF, const float* measures,
                              const float* sortedPos, const float* sortedVel, const uint32_t* index,
                              const uint32_t* cellStart, const uint32_t* cellEnd, uint32_t n,
                              hipStream_t s) {
    if (n == 0) return CMAG_OK;
    go_peri<true, true>(P, viscF, presF, measures, sortedPos, sortedVel, index, cellStart, cellEnd, n, s);
    return check_launch("peristalsis_force_kernel<fused>");
}

}  // namespace cmag

using namespace cmag;

extern "C" {

int cmag_density(const cmag_params* p, float* measures, const float* sortedPos,
                 const uint32_t* cellStart, const uint32_t* cellEnd, uint32_t n,
                 cmag_stream_t stream) {
    CMAG_TRY(validate_params(p));
    if (n && (!measures || !sortedPos || !cellStart || !cellEnd))
        return fail(CMAG_E_INVALID, "cmag_density: null pointer");
    return launch_density(*p, measures, sortedPos, cellStart, cellEnd, n, as_stream(stream));
}

int cmag_force(const cmag_params* p, float* acc, const float* measures, const float* sortedPos,
               const float* sortedVel, const uint32_t* index, const uint32_t* cellStart,
               const uint32_t* cellEnd, uint32_t n, cmag_stream_t stream) {
    CMAG_TRY(validate_params(p));
    if (n && (!acc || !measures || !sortedPos || !sortedVel || !index || !cellStart || !cellEnd))
        return fail(CMAG_E_INVALID, "cmag_force: null pointer");
    return launch_force(*p, acc, measures, sortedPos, sortedVel, index, cellStart, cellEnd, n,
                        as_stream(stream));
}

int cmag_pressure_force(const cmag_params* p, float* presF, const float* measures,
                        const float* sortedPos, const uint32_t* index, const uint32_t* cellStart,
                        const uint32_t* cellEnd, uint32_t n, cmag_stream_t stream) {
    CMAG_TRY(validate_params(p));
    if (n && (!presF || !measures || !sortedPos || !index || !cellStart || !cellEnd))
        return fail(CMAG_E_INVALID, "cmag_pressure_force: null pointer");
    return launch_pressure_force(*p, presF, measures, sortedPos, index, cellStart, cellEnd, n,
                                 as_stream(stream));
}

int cmag_viscous_force(const cmag_params* p, float* viscF, const float* measures,
                       const float* sortedPos, const float* sortedVel, const uint32_t* index,
                       const uint32_t* cellStart, const uint32_t* cellEnd, uint32_t n,
                       cmag_stream_t stream) {
    CMAG_TRY(validate_params(p));
    if (n && (!viscF || !measures || !sortedPos || !sortedVel || !index || !cellStart || !cellEnd))
        return fail(CMAG_E_INVALID, "cmag_viscous_force: null pointer");
    return launch_viscous_force(*p, viscF, measures, sortedPos, sortedVel, index, cellStart, cellEnd,
                                n, as_stream(stream));
}

}  // extern "C"
