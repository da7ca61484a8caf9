// adjoint.hip - hand-written vector-Jacobian products of the Poiseuille step
// (gfx950, FP32 with FP64 parameter accumulators).
//
// The reference has no adjoint.  These differentiate this library's forward
// passes with the neighbour structure (sort order, cell lists) of the step
// held fixed, as it is piecewise constant:
//   density_vjp    density + EOS      poiseuilleFlowKernel.cu:99-182
//   force_vjp      pressure + Morris viscous force with the wall
//                  extrapolation      poiseuilleFlowKernel.cu:184-309
//   integrate_vjp  integratePoiseuilleSystemD  poiseuilleFlowKernel.cu:311-348
// Walls (w == 1) are not state: their density is the constant rho0 and their
// adjoints are zero.
//
// Pair notation (target k, neighbour j, r = x_k - x_j after the image shift,
// d = |r|, G = (dW/dr)/d, G' = dG/dd, pi = p/rho^2):
//   a_k = sum_j [ -m (pi_k + pi_j) G r + c_kj V_kj ],  c_kj = 2 m mu G/(rho_k rho_j),
//   V_kj = l_k - l_j, or beta l_k for a wall neighbour beyond its wall plane.
// A pair enters k's adjoint once as target and once as neighbour; with
// D = abar_k - abar_j (abar of walls is zero) both collapse into one term, so
// every pass gathers over the same stencil as the forward pass, with no
// atomics:
//   pibar_k  = sum -m G (r . D)
//   rhobar_k = pibar_k dpi/drho - (1/rho_k) sum c (V . D)
//   lbar_k   = sum c (beta~ abar_k - abar_j)        (beta~ = beta or 1)
//   xbar_k   = sum A [G D + (G'/d) r (r . D)] + B (G'/d) r (V . D)
//              (+ c (abar_k . l_k) dbeta/dy_k for wall neighbours)
//   mubar    = sum_k sum_j 2 m G/(rho_k rho_j) (abar_k . V_kj)
//   density: xbar_k += sum_j m G r (rhobar_k + rhobar_j)
#include <string>
#include <type_traits>

#include "kernels.hpp"
#include "sph_device.hpp"

namespace cmag {

// d/dq of the dW/dr bracket: -4 t^3 + 1.5 t^2 (2q + 1), t = 1 - q/2
__device__ __forceinline__ float grad_bracket_dq(float q) {
    const float t = 1.0f - 0.5f * q;
    return -4.0f * t * t * t + 1.5f * t * t * (2.0f * q + 1.0f);
}

__device__ __forceinline__ float grad_bracket(float q) {
    const float t = 1.0f - 0.5f * q;
    const float t2 = t * t;
    return t2 * t2 - t2 * t * (2.0f * q + 1.0f);
}

// d p / d rho of the equation of state
__device__ __forceinline__ float deos(const cmag_params& P, float dens) {
    if (P.eos == CMAG_EOS_TAIT)
        return P.B * P.gamma * pow_int(dens / P.rest_density, (int)P.gamma - 1) / P.rest_density;
    if (P.eos == CMAG_EOS_LINEAR_C2) return P.soundspeed * P.soundspeed;
    return 50.0f * P.soundspeed;
}

__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    return ax * bx + ay * by + az * bz;
}

// ---- integrate ----------------------------------------------------------------
// unsorted order; gpos/gvel/gvelLF: adjoints of the step outputs in, of its
// inputs out (velLF is no input of integrate); gacc out; ggrav[i] += share of
// the body-force adjoint
__global__ __launch_bounds__(256) void integrate_vjp_kernel(const cmag_params P, const float4* __restrict__ pos,
                                                            float4* __restrict__ gpos, float4* __restrict__ gvel,
                                                            float4* __restrict__ gvelLF, float4* __restrict__ gacc,
                                                            double* __restrict__ ggrav, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
    if (pos[i].w == CMAG_PO_WALL) {
        gpos[i] = z;
        gvel[i] = z;
        gvelLF[i] = z;
        gacc[i] = z;
        return;
    }
    const float dt = P.delta_time;
    const float4 gx = gpos[i], gv = gvel[i], gl = gvelLF[i];
    const float ux = gv.x + 0.5f * gl.x + dt * gx.x;
    const float uy = gv.y + 0.5f * gl.y + dt * gx.y;
    const float uz = gv.z + 0.5f * gl.z + dt * gx.z;
    gacc[i] = make_float4(dt * ux, dt * uy, dt * uz, 0.f);
    gvel[i] = make_float4(ux + 0.5f * gl.x, uy + 0.5f * gl.y, uz + 0.5f * gl.z, 0.f);
    gvelLF[i] = z;
    if (ggrav) {
        ggrav[4 * (size_t)i] += (double)(dt * ux);
        ggrav[4 * (size_t)i + 1] += (double)(dt * uy);
        ggrav[4 * (size_t)i + 2] += (double)(dt * uz);
    }
}

__global__ __launch_bounds__(256) void gather_f4_kernel(float4* __restrict__ dst, const float4* __restrict__ src,
                                                        const uint32_t* __restrict__ index, uint32_t n) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t i = index[k];
    dst[k] = i < n ? src[i] : make_float4(0.f, 0.f, 0.f, 0.f);
}

// ---- force ----------------------------------------------------------------------
// sorted order.  gmu_acc (optional) += the target's share of d/d mu in FP64;
// gmu (optional) = the same share in FP32
template <int DIM, int HASH, typename Cells>
__global__ __launch_bounds__(256) void force_vjp_kernel(const cmag_params P, float4* __restrict__ gpos,
                                                        float4* __restrict__ gvel, float* __restrict__ grho,
                                                        float* __restrict__ gmu, double* __restrict__ gmu_acc,
                                                        const float4* __restrict__ gacc,
                                                        const float4* __restrict__ measures,
                                                        const float4* __restrict__ sortedPos,
                                                        const float4* __restrict__ sortedVel, const Cells cells,
                                                        uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 pi = sortedPos[i];
    const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
    if (pi.w == CMAG_PO_WALL) {
        gpos[i] = z;
        gvel[i] = z;
        grho[i] = 0.f;
        if (gmu) gmu[i] = 0.f;
        return;
    }
    const float4 li = sortedVel[i], ai = gacc[i], mi = measures[i];
    const float rho_i = mi.x, pi_i = mi.y / (mi.x * mi.x);
    const float m = P.particle_mass, h = P.smoothing_radius, gc = P.grad_coeff;
    const float r2max = 4.0f * h * h;
    const float top = P.wall_top, bottom = P.wall_bottom;
    float pibar = 0.f, rhobar = 0.f, mubar = 0.f;
    float gx = 0.f, gy = 0.f, gz = 0.f, lx = 0.f, ly = 0.f, lz = 0.f;
    const Cell gp = grid_pos(P, pi.x, pi.y, pi.z);
    walk_stencil<DIM, HASH>(P, pi, gp, cells, [&](uint32_t j, float qx, float qy, float qz, int) {
        if (j == i) return;
        const float4 pj = sortedPos[j];
        const float rx = qx - pj.x, ry = qy - pj.y, rz = qz - pj.z;
        const float r2 = rx * rx + ry * ry + rz * rz;
        if (!(r2 < r2max)) return;
        const float dist = sqrtf(r2);
        const float q = dist / h;
        const float4 lj = sortedVel[j], aj = gacc[j], mj = measures[j];
        const float rho_j = mj.x, pi_j = mj.y / (mj.x * mj.x);
        const float temp = gc * grad_bracket(q);
        const float dtemp = gc / h * grad_bracket_dq(q);
        const float G = temp / dist;
        const float Gp = (dtemp * dist - temp) / r2;
        float bt = 1.f, dbeta = 0.f;
        bool wall_beyond = false;
        if (pj.w == CMAG_PO_WALL && pj.y > top) {
            const float dA = top - pi.y, dB = pj.y - top;
            const float b = 1.f + dB / dA;
            wall_beyond = true;
            bt = fminf(P.beta_max, b);
            if (b < P.beta_max) dbeta = dB / (dA * dA);
        } else if (pj.w == CMAG_PO_WALL && pj.y < bottom) {
            const float dA = pi.y - bottom, dB = bottom - pj.y;
            const float b = 1.f + dB / dA;
            wall_beyond = true;
            bt = fminf(P.beta_max, b);
            if (b < P.beta_max) dbeta = -dB / (dA * dA);
        }
        float Vx, Vy, Vz;
        if (wall_beyond) {
            Vx = bt * li.x; Vy = bt * li.y; Vz = bt * li.z;
        } else {
            Vx = li.x - lj.x; Vy = li.y - lj.y; Vz = li.z - lj.z;
        }
        const float Dx = ai.x - aj.x, Dy = ai.y - aj.y, Dz = ai.z - aj.z;
        const float A = -m * (pi_i + pi_j);
        const float B = 2.f * m * P.mu / (rho_i * rho_j);
        const float c = B * G;
        const float rD = dot3(rx, ry, rz, Dx, Dy, Dz);
        const float VD = dot3(Vx, Vy, Vz, Dx, Dy, Dz);
        pibar += -m * G * rD;
        rhobar += -(c / rho_i) * VD;
        mubar += 2.f * m * G / (rho_i * rho_j) * dot3(ai.x, ai.y, ai.z, Vx, Vy, Vz);
        const float lt = wall_beyond ? bt : 1.f;
        lx += c * (lt * ai.x - aj.x);
        ly += c * (lt * ai.y - aj.y);
        lz += c * (lt * ai.z - aj.z);
        const float k1 = A * G, k2 = (A * rD + B * VD) * Gp / dist;
        gx += k1 * Dx + k2 * rx;
        gy += k1 * Dy + k2 * ry;
        gz += k1 * Dz + k2 * rz;
        if (wall_beyond) gy += c * dot3(ai.x, ai.y, ai.z, li.x, li.y, li.z) * dbeta;
    });
    const float dpi = deos(P, rho_i) / (rho_i * rho_i) - 2.f * mi.y / (rho_i * rho_i * rho_i);
    gpos[i] = make_float4(gx, gy, gz, 0.f);
    gvel[i] = make_float4(lx, ly, lz, 0.f);
    grho[i] = pibar * dpi + rhobar;
    if (gmu) gmu[i] = mubar;
    if (gmu_acc) gmu_acc[i] += (double)mubar;
}

// ---- density ------------------------------------------------------------------
// sorted order: gposS[k] += sum_j m G r (grho_k + grho_j) for non-walls.  With
// index, the total goes to the unsorted adjoints: gpos[index[k]] += gposS[k],
// gvelLF[index[k]] = gvelS[k] (the step's output adjoints already consumed)
template <int DIM, int HASH, typename Cells>
__global__ __launch_bounds__(256) void density_vjp_kernel(const cmag_params P, float4* __restrict__ gposS,
                                                          const float* __restrict__ grho,
                                                          const float4* __restrict__ sortedPos, const Cells cells,
                                                          uint32_t n, const uint32_t* __restrict__ index,
                                                          const float4* __restrict__ gvelS,
                                                          float4* __restrict__ gpos, float4* __restrict__ gvelLF) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 pi = sortedPos[i];
    if (pi.w == CMAG_PO_WALL) return;
    const float m = P.particle_mass, h = P.smoothing_radius, gc = P.grad_coeff;
    const float r2max = 4.0f * h * h;
    const float gi = grho[i];
    float gx = 0.f, gy = 0.f, gz = 0.f;
    const Cell gp = grid_pos(P, pi.x, pi.y, pi.z);
    walk_stencil<DIM, HASH>(P, pi, gp, cells, [&](uint32_t j, float qx, float qy, float qz, int) {
        if (j == i) return;
        const float4 pj = sortedPos[j];
        const float rx = qx - pj.x, ry = qy - pj.y, rz = qz - pj.z;
        const float r2 = rx * rx + ry * ry + rz * rz;
        if (!(r2 < r2max)) return;
        const float dist = sqrtf(r2);
        const float G = gc * grad_bracket(dist / h) / dist;
        const float c = m * G * (gi + grho[j]);
        gx += c * rx; gy += c * ry; gz += c * rz;
    });
    float4 g = gposS[i];
    g.x += gx; g.y += gy; g.z += gz;
    if (!index) {
        gposS[i] = g;
        return;
    }
    const uint32_t k = index[i];
    if (k >= n) return;
    float4 t = gpos[k];
    gpos[k] = make_float4(t.x + g.x, t.y + g.y, t.z + g.z, t.w);
    gvelLF[k] = gvelS[i];
}

// ---- launchers ------------------------------------------------------------------
template <typename F>
static int dispatch_dh(const cmag_params& P, F&& f) {
    const bool mo = P.hash_order == CMAG_HASH_MORTON;
    if (P.dim == 3) {
        if (mo) f(std::integral_constant<int, 3>{}, std::integral_constant<int, 1>{});
        else f(std::integral_constant<int, 3>{}, std::integral_constant<int, 0>{});
    } else {
        if (mo) f(std::integral_constant<int, 2>{}, std::integral_constant<int, 1>{});
        else f(std::integral_constant<int, 2>{}, std::integral_constant<int, 0>{});
    }
    return CMAG_OK;
}

static int need_poiseuille(const cmag_params& P, const char* who) {
    if (P.system != CMAG_SYS_POISEUILLE)
        return fail(CMAG_E_INVALID, std::string(who) + ": the adjoint covers the Poiseuille system");
    return CMAG_OK;
}

int launch_integrate_vjp(const cmag_params& P, const float* pos, float* gpos, float* gvel, float* gvelLF,
                         float* gacc, double* ggrav, uint32_t n, hipStream_t s) {
    CMAG_TRY(need_poiseuille(P, "integrate_vjp"));
    if (n == 0) return CMAG_OK;
    hipLaunchKernelGGL(integrate_vjp_kernel, dim3(div_up(n, 256)), dim3(256), 0, s, P,
                       reinterpret_cast<const float4*>(pos), reinterpret_cast<float4*>(gpos),
                       reinterpret_cast<float4*>(gvel), reinterpret_cast<float4*>(gvelLF),
                       reinterpret_cast<float4*>(gacc), ggrav, n);
    return check_launch("integrate_vjp_kernel");
}

int launch_gather_f4(float* dst, const float* src, const uint32_t* index, uint32_t n, hipStream_t s) {
    if (n == 0) return CMAG_OK;
    hipLaunchKernelGGL(gather_f4_kernel, dim3(div_up(n, 256)), dim3(256), 0, s, reinterpret_cast<float4*>(dst),
                       reinterpret_cast<const float4*>(src), index, n);
    return check_launch("gather_f4_kernel");
}

int launch_force_vjp(const cmag_params& P, float* gposS, float* gvelS, float* grho, float* gmu, double* gmu_acc,
                     const float* gaccS, const float* measures, const float* sortedPos, const float* sortedVel,
                     const uint32_t* cellStart, const uint32_t* cellEnd, const uint32_t* cellFirst, uint32_t n,
                     hipStream_t s) {
    CMAG_TRY(need_poiseuille(P, "force_vjp"));
    if (n == 0) return CMAG_OK;
    auto f4 = [](float* q) { return reinterpret_cast<float4*>(q); };
    auto c4 = [](const float* q) { return reinterpret_cast<const float4*>(q); };
    dispatch_dh(P, [&](auto D, auto H) {
        constexpr int DIM = decltype(D)::value, HASH = decltype(H)::value;
        if (cellFirst)
            hipLaunchKernelGGL((force_vjp_kernel<DIM, HASH, FirstCells>), dim3(div_up(n, 256)), dim3(256), 0, s,
                               P, f4(gposS), f4(gvelS), grho, gmu, gmu_acc, c4(gaccS), c4(measures), c4(sortedPos),
                               c4(sortedVel), FirstCells{cellFirst}, n);
        else
            hipLaunchKernelGGL((force_vjp_kernel<DIM, HASH, RefCells>), dim3(div_up(n, 256)), dim3(256), 0, s, P,
                               f4(gposS), f4(gvelS), grho, gmu, gmu_acc, c4(gaccS), c4(measures), c4(sortedPos),
                               c4(sortedVel), RefCells{cellStart, cellEnd, n}, n);
    });
    return check_launch("force_vjp_kernel");
}

int launch_density_vjp(const cmag_params& P, float* gposS, const float* grho, const float* sortedPos,
                       const uint32_t* cellStart, const uint32_t* cellEnd, const uint32_t* cellFirst, uint32_t n,
                       const uint32_t* index, const float* gvelS, float* gpos, float* gvelLF, hipStream_t s) {
    CMAG_TRY(need_poiseuille(P, "density_vjp"));
    if (n == 0) return CMAG_OK;
    auto f4 = [](float* q) { return reinterpret_cast<float4*>(q); };
    auto c4 = [](const float* q) { return reinterpret_cast<const float4*>(q); };
    dispatch_dh(P, [&](auto D, auto H) {
        constexpr int DIM = decltype(D)::value, HASH = decltype(H)::value;
        if (cellFirst)
            hipLaunchKernelGGL((density_vjp_kernel<DIM, HASH, FirstCells>), dim3(div_up(n, 256)), dim3(256), 0, s,
                               P, f4(gposS), grho, c4(sortedPos), FirstCells{cellFirst}, n, index, c4(gvelS),
                               f4(gpos), f4(gvelLF));
        else
            hipLaunchKernelGGL((density_vjp_kernel<DIM, HASH, RefCells>), dim3(div_up(n, 256)), dim3(256), 0, s, P,
                               f4(gposS), grho, c4(sortedPos), RefCells{cellStart, cellEnd, n}, n, index,
                               c4(gvelS), f4(gpos), f4(gvelLF));
    });
    return check_launch("density_vjp_kernel");
}

}  // namespace cmag

using namespace cmag;

extern "C" {

int cmag_integrate_vjp(const cmag_params* p, const float* pos, float* gpos, float* gvel, float* gvelLF,
                       float* gacc, double* ggrav, uint32_t n, cmag_stream_t stream) {
    CMAG_TRY(validate_params(p));
    if (n && (!pos || !gpos || !gvel || !gvelLF || !gacc))
        return fail(CMAG_E_INVALID, "cmag_integrate_vjp: null pointer");
    return launch_integrate_vjp(*p, pos, gpos, gvel, gvelLF, gacc, ggrav, n, as_stream(stream));
}

int cmag_force_vjp(const cmag_params* p, float* gposS, float* gvelS, float* grho, float* gmu, const float* gaccS,
                   const float* measures, const float* sortedPos, const float* sortedVel,
                   const uint32_t* cellStart, const uint32_t* cellEnd, uint32_t n, cmag_stream_t stream) {
    CMAG_TRY(validate_params(p));
    if (n && (!gposS || !gvelS || !grho || !gaccS || !measures || !sortedPos || !sortedVel || !cellStart ||
              !cellEnd))
        return fail(CMAG_E_INVALID, "cmag_force_vjp: null pointer");
    return launch_force_vjp(*p, gposS, gvelS, grho, gmu, nullptr, gaccS, measures, sortedPos, sortedVel, cellStart,
                            cellEnd, nullptr, n, as_stream(stream));
}

int cmag_density_vjp(const cmag_params* p, float* gposS, const float* grho, const float* sortedPos,
                     const uint32_t* cellStart, const uint32_t* cellEnd, uint32_t n, cmag_stream_t stream) {
    CMAG_TRY(validate_params(p));
    if (n && (!gposS || !grho || !sortedPos || !cellStart || !cellEnd))
        return fail(CMAG_E_INVALID, "cmag_density_vjp: null pointer");
    return launch_density_vjp(*p, gposS, grho, sortedPos, cellStart, cellEnd, nullptr, n, nullptr, nullptr, nullptr,
                              nullptr, as_stream(stream));
}

}  // extern "C"
