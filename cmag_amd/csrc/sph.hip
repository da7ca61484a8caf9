// sph.hip - neighbour-pass kernels of one system and their launchers.
// The per-particle bodies live in sph_device.hpp (shared with batch.hip).
#include "kernels.hpp"
#include "sph_device.hpp"

namespace cmag {

template <int SYS, int DIM, int HASH, typename Cells>
__global__ __launch_bounds__(256) void density_kernel(const cmag_params P, float4* __restrict__ measures,
                                                       const float4* __restrict__ sortedPos, const Cells cells,
                                                       uint32_t n) {
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= n) return;
    density_body<SYS, DIM, HASH>(P, i, measures, sortedPos, cells);
}

template <int SYS, int DIM, int HASH, typename Cells>
__global__ __launch_bounds__(256) void force_kernel(const cmag_params P, float4* __restrict__ acc,
                                                     const float4* __restrict__ measures,
                                                     const float4* __restrict__ sortedPos,
                                                     const float4* __restrict__ sortedVel,
                                                     const uint32_t* __restrict__ index, const Cells cells,
                                                     uint32_t n) {
    __shared__ uint32_t nbl[256 / 64][NB_CAP][64];
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= n) return;
    force_body<SYS, DIM, HASH>(P, i, acc, measures, sortedPos, sortedVel, index, cells, n, nbl[threadIdx.x >> 6]);
}

// Peristalsis: pressure and viscous passes; the fused variant produces both
// from one neighbour walk (same per-pair arithmetic as the split kernels)
template <int DIM, int HASH, bool WANT_VISC, bool WANT_PRES, typename Cells>
__global__ __launch_bounds__(256) void peristalsis_force_kernel(
    const cmag_params P, float4* __restrict__ viscF, float4* __restrict__ presF,
    const float4* __restrict__ measures, const float4* __restrict__ sortedPos,
    const float4* __restrict__ sortedVel, const uint32_t* __restrict__ index, const Cells cells,
    uint32_t n) {
    __shared__ uint32_t nbl[256 / 64][NB_CAP][64];
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= n) return;
    peristalsis_force_body<DIM, HASH, WANT_VISC, WANT_PRES>(P, i, viscF, presF, measures, sortedPos, sortedVel,
                                                            index, cells, n, nbl[threadIdx.x >> 6]);
}

// in-range neighbours (j != i, r^2 < 4h^2) of each sorted particle: the pair
// count behind the neighbour-pair interactions/s figure
template <int SYS, int DIM, int HASH, typename Cells>
__global__ __launch_bounds__(256) void count_kernel(const cmag_params P, uint32_t* __restrict__ counts,
                                                     const float4* __restrict__ sortedPos, const Cells cells,
                                                     uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 pi = sortedPos[i];
    const Cell gp = grid_pos(P, pi.x, pi.y, pi.z);
    const float r2max = 4.0f * P.smoothing_radius * P.smoothing_radius;
    uint32_t c = 0;
    walk_stencil<DIM, HASH>(P, pi, gp, cells, [&](uint32_t j, float px, float py, float pz, int) {
        const float4 pj = sortedPos[j];
        const float rx = px - pj.x, ry = py - pj.y, rz = pz - pj.z;
        c += (j != i && norm2(rx, ry, rz) < r2max) ? 1u : 0u;
    });
    counts[i] = c;
}

// ---- launchers --------------------------------------------------------------
// every kernel is instantiated per (system, dimension, hash order, range policy)
template <template <int, int, int> class K, typename... Args>
static void dispatch3(const cmag_params& P, Args&&... args) {
    const bool m = P.hash_order == CMAG_HASH_MORTON;
#define CMAG_GO(SYS, DIM, H) K<SYS, DIM, H>::go(args...)
    if (P.dim == 3) {
        if (P.system == CMAG_SYS_DAMBREAK) { if (m) CMAG_GO(0, 3, 1); else CMAG_GO(0, 3, 0); }
        else if (P.system == CMAG_SYS_POISEUILLE) { if (m) CMAG_GO(1, 3, 1); else CMAG_GO(1, 3, 0); }
        else { if (m) CMAG_GO(2, 3, 1); else CMAG_GO(2, 3, 0); }
    } else {
        if (P.system == CMAG_SYS_DAMBREAK) { if (m) CMAG_GO(0, 2, 1); else CMAG_GO(0, 2, 0); }
        else if (P.system == CMAG_SYS_POISEUILLE) { if (m) CMAG_GO(1, 2, 1); else CMAG_GO(1, 2, 0); }
        else { if (m) CMAG_GO(2, 2, 1); else CMAG_GO(2, 2, 0); }
    }
#undef CMAG_GO
}

template <int SYS, int DIM, int H>
struct DensityGo {
    template <typename Cells>
    static void run(const cmag_params& P, float* measures, const float* sortedPos, const Cells& cells,
                    uint32_t n, hipStream_t s) {
        hipLaunchKernelGGL((density_kernel<SYS, DIM, H, Cells>), dim3(div_up(n, 256)), dim3(256), 0, s, P,
                           reinterpret_cast<float4*>(measures), reinterpret_cast<const float4*>(sortedPos),
                           cells, n);
    }
    static void go(const cmag_params& P, float* measures, const float* sortedPos, const uint32_t* cellStart,
                   const uint32_t* cellEnd, const uint32_t* cellFirst, uint32_t n, hipStream_t s) {
        if (cellFirst) run(P, measures, sortedPos, FirstCells{cellFirst}, n, s);
        else run(P, measures, sortedPos, RefCells{cellStart, cellEnd, n}, n, s);
    }
};

int launch_density(const cmag_params& P, float* measures, const float* sortedPos,
                   const uint32_t* cellStart, const uint32_t* cellEnd, uint32_t n, hipStream_t s,
                   const uint32_t* cellFirst) {
    if (n == 0) return CMAG_OK;
    dispatch3<DensityGo>(P, P, measures, sortedPos, cellStart, cellEnd, cellFirst, n, s);
    return check_launch("density_kernel");
}

template <int SYS, int DIM, int H>
struct CountGo {
    static void go(const cmag_params& P, uint32_t* counts, const float* sortedPos, const uint32_t* cellStart,
                   const uint32_t* cellEnd, uint32_t n, hipStream_t s) {
        hipLaunchKernelGGL((count_kernel<SYS, DIM, H, RefCells>), dim3(div_up(n, 256)), dim3(256), 0, s, P,
                           counts, reinterpret_cast<const float4*>(sortedPos), RefCells{cellStart, cellEnd, n}, n);
    }
};

template <int SYS, int DIM, int H>
struct ForceGo {
    template <typename Cells>
    static void run(const cmag_params& P, float* acc, const float* measures, const float* sortedPos,
                    const float* sortedVel, const uint32_t* index, const Cells& cells, uint32_t n,
                    hipStream_t s) {
        if constexpr (SYS != CMAG_SYS_PERISTALSIS)
            hipLaunchKernelGGL((force_kernel<SYS, DIM, H, Cells>), dim3(div_up(n, 256)), dim3(256), 0, s, P,
                               reinterpret_cast<float4*>(acc), reinterpret_cast<const float4*>(measures),
                               reinterpret_cast<const float4*>(sortedPos),
                               reinterpret_cast<const float4*>(sortedVel), index, cells, n);
    }
    static void go(const cmag_params& P, float* acc, const float* measures, const float* sortedPos,
                   const float* sortedVel, const uint32_t* index, const uint32_t* cellStart,
                   const uint32_t* cellEnd, const uint32_t* cellFirst, uint32_t n, hipStream_t s) {
        if (cellFirst) run(P, acc, measures, sortedPos, sortedVel, index, FirstCells{cellFirst}, n, s);
        else run(P, acc, measures, sortedPos, sortedVel, index, RefCells{cellStart, cellEnd, n}, n, s);
    }
};

int launch_force(const cmag_params& P, float* acc, const float* measures, const float* sortedPos,
                 const float* sortedVel, const uint32_t* index, const uint32_t* cellStart,
                 const uint32_t* cellEnd, uint32_t n, hipStream_t s, const uint32_t* cellFirst) {
    if (n == 0) return CMAG_OK;
    if (n >= (1u << NB_SHIFT_BITS)) return fail(CMAG_E_INVALID, "force passes take at most 2^27 particles");
    if (P.system == CMAG_SYS_PERISTALSIS)
        return fail(CMAG_E_INVALID, "Peristalsis uses the split pressure/viscous passes");
    dispatch3<ForceGo>(P, P, acc, measures, sortedPos, sortedVel, index, cellStart, cellEnd, cellFirst, n, s);
    return check_launch("force_kernel");
}

template <bool V, bool PR, typename Cells>
static void go_peri_cells(const cmag_params& P, float* viscF, float* presF, const float* measures,
                          const float* sortedPos, const float* sortedVel, const uint32_t* index,
                          const Cells& cells, uint32_t n, hipStream_t s) {
    const bool m = P.hash_order == CMAG_HASH_MORTON;
    auto launch = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(div_up(n, 256)), dim3(256), 0, s, P, reinterpret_cast<float4*>(viscF),
                           reinterpret_cast<float4*>(presF), reinterpret_cast<const float4*>(measures),
                           reinterpret_cast<const float4*>(sortedPos), reinterpret_cast<const float4*>(sortedVel),
                           index, cells, n);
    };
    if (P.dim == 3) {
        if (m) launch(peristalsis_force_kernel<3, 1, V, PR, Cells>);
        else launch(peristalsis_force_kernel<3, 0, V, PR, Cells>);
    } else {
        if (m) launch(peristalsis_force_kernel<2, 1, V, PR, Cells>);
        else launch(peristalsis_force_kernel<2, 0, V, PR, Cells>);
    }
}

template <bool V, bool PR>
static void go_peri(const cmag_params& P, float* viscF, float* presF, const float* measures,
                    const float* sortedPos, const float* sortedVel, const uint32_t* index,
                    const uint32_t* cellStart, const uint32_t* cellEnd, const uint32_t* cellFirst, uint32_t n,
                    hipStream_t s) {
    if (n >= (1u << NB_SHIFT_BITS)) return;  // rejected by the callers
    if (cellFirst)
        go_peri_cells<V, PR>(P, viscF, presF, measures, sortedPos, sortedVel, index, FirstCells{cellFirst}, n, s);
    else
        go_peri_cells<V, PR>(P, viscF, presF, measures, sortedPos, sortedVel, index,
                             RefCells{cellStart, cellEnd, n}, n, s);
}

int launch_pressure_force(const cmag_params& P, float* presF, const float* measures,
                          const float* sortedPos, const uint32_t* index, const uint32_t* cellStart,
                          const uint32_t* cellEnd, uint32_t n, hipStream_t s) {
    if (n == 0) return CMAG_OK;
    if (n >= (1u << NB_SHIFT_BITS)) return fail(CMAG_E_INVALID, "force passes take at most 2^27 particles");
    go_peri<false, true>(P, nullptr, presF, measures, sortedPos, nullptr, index, cellStart, cellEnd, nullptr, n, s);
    return check_launch("peristalsis_force_kernel<pressure>");
}

int launch_viscous_force(const cmag_params& P, float* viscF, const float* measures,
                         const float* sortedPos, const float* sortedVel, const uint32_t* index,
                         const uint32_t* cellStart, const uint32_t* cellEnd, uint32_t n,
                         hipStream_t s) {
    if (n == 0) return CMAG_OK;
    if (n >= (1u << NB_SHIFT_BITS)) return fail(CMAG_E_INVALID, "force passes take at most 2^27 particles");
    go_peri<true, false>(P, viscF, nullptr, measures, sortedPos, sortedVel, index, cellStart, cellEnd, nullptr, n, s);
    return check_launch("peristalsis_force_kernel<viscous>");
}

int launch_peristalsis_forces(const cmag_params& P, float* viscF, float* presF, const float* measures,
                              const float* sortedPos, const float* sortedVel, const uint32_t* index,
                              const uint32_t* cellStart, const uint32_t* cellEnd, uint32_t n,
                              hipStream_t s, const uint32_t* cellFirst) {
    if (n == 0) return CMAG_OK;
    if (n >= (1u << NB_SHIFT_BITS)) return fail(CMAG_E_INVALID, "force passes take at most 2^27 particles");
    go_peri<true, true>(P, viscF, presF, measures, sortedPos, sortedVel, index, cellStart, cellEnd, cellFirst, n, s);
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

int cmag_count_neighbours(const cmag_params* p, uint32_t* counts, const float* sortedPos,
                          const uint32_t* cellStart, const uint32_t* cellEnd, uint32_t n,
                          cmag_stream_t stream) {
    CMAG_TRY(validate_params(p));
    if (n && (!counts || !sortedPos || !cellStart || !cellEnd))
        return fail(CMAG_E_INVALID, "cmag_count_neighbours: null pointer");
    if (n == 0) return CMAG_OK;
    dispatch3<CountGo>(*p, *p, counts, sortedPos, cellStart, cellEnd, n, as_stream(stream));
    return check_launch("count_kernel");
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
