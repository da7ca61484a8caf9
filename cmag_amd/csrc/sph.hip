// sph.hip - neighbour-pass kernels of one system and their launchers.
// The per-particle bodies live in sph_device.hpp (shared with batch.hip).
#include <type_traits>

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

// ---- sub-cell walk order (DamBreak) ---------------------------------------------
// The bucket of every sorted row (a byte: fluid eighth 0..7 of the cell
// along x, walls 8) comes from the reorder.  subcell_move_kernel: row i of
// the reference order (cell c = hash[i]) moves to row first[c] + below of the
// sub-cell order, below = the rows of the cell in a lower bucket or in the
// same bucket and earlier in reference order; the cell's first row also
// writes the running bucket counts pre[c] (eight 16-bit counts; cells without
// particles keep the zeros they were cleared to).  O(cell size) work per row (~18 at the
// measured sizes) over 16-byte loads of the buckets.
__global__ __launch_bounds__(256) void subcell_move_kernel(const uint32_t* __restrict__ hash,
                                                           const uint32_t* __restrict__ index,
                                                           const float4* __restrict__ sortedPos,
                                                           const float4* __restrict__ sortedVel,
                                                           const uint32_t* __restrict__ first,
                                                           const uint8_t* __restrict__ bucket, uint32_t n,
                                                           uint4* __restrict__ pre, float4* __restrict__ posW,
                                                           float4* __restrict__ velW, uint32_t* __restrict__ indexW,
                                                           uint32_t* __restrict__ w2r) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = hash[i];
    const uint32_t a = first[c], b = min(first[c + 1], n);
    const uint32_t bi = bucket[i];
    uint32_t below = 0;
    uint32_t cnt[X_SUBS] = {};  // fluid rows of eighths <= t
    // the cell's bucket bytes, 16 per load from a 16-byte boundary (the array
    // is padded so a load never reads past its end)
    for (uint32_t k16 = a & ~15u; k16 < b; k16 += 16u) {
        const uint4 v = *reinterpret_cast<const uint4*>(bucket + k16);
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
        // byte t of acc counts this chunk's rows of eighths <= t (at most 16)
        uint64_t acc = 0;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const uint32_t k = k16 + (uint32_t)j;
            if (k < a || k >= b) continue;
            const uint32_t bk = (w[j >> 2] >> (8 * (j & 3))) & 0xffu;
            below += (bk < bi || (bk == bi && k < i)) ? 1u : 0u;
            acc += bk < (uint32_t)X_SUBS ? 0x0101010101010101ull << (8u * bk) : 0ull;
        }
#pragma unroll
        for (int t = 0; t < X_SUBS; ++t) cnt[t] += (uint32_t)(acc >> (8 * t)) & 0xffu;
    }
    const uint32_t w = a + below;
    if (w >= n) return;  // inconsistent cell arrays only
    // the walk tells walls from fluid by the sign bit of w (the density pass
    // stores -rho in fluid rows): every non-fluid row, whatever its type code
    // (-0.0 and negative codes included), gets the non-negative tag +0
    float4 p = sortedPos[i];
    if (p.w != CMAG_DB_FLUID) p.w = 0.0f;
    posW[w] = p;
    velW[w] = sortedVel[i];
    indexW[w] = index[i];
    w2r[w] = i;
    // 16-bit counts, capped: a cell of more than 65,535 rows (no physical
    // state has one) keeps its runs inside the cell
    if (i == a) {
        uint32_t h[X_SUBS];
#pragma unroll
        for (int t = 0; t < X_SUBS; ++t) h[t] = min(cnt[t], 0xffffu);
        pre[c] = make_uint4(h[0] | h[1] << 16, h[2] | h[3] << 16, h[4] | h[5] << 16, h[6] | h[7] << 16);
    }
}

int launch_subcell_build(const cmag_params& P, const uint32_t* hash, const uint32_t* index, const float* sortedPos,
                         const float* sortedVel, const uint32_t* cellFirst, uint32_t n, uint32_t numCells,
                         uint32_t* pre, const uint8_t* bucket, float* posW, float* velW, uint32_t* indexW,
                         uint32_t* w2r, hipStream_t s) {
    (void)P;
    if (n == 0) return CMAG_OK;
    CMAG_TRY(launch_fill_u32(pre, 0u, 4 * (size_t)numCells, s));
    hipLaunchKernelGGL(subcell_move_kernel, dim3(div_up(n, 256)), dim3(256), 0, s, hash, index,
                       reinterpret_cast<const float4*>(sortedPos), reinterpret_cast<const float4*>(sortedVel),
                       cellFirst, bucket, n, reinterpret_cast<uint4*>(pre), reinterpret_cast<float4*>(posW),
                       reinterpret_cast<float4*>(velW), indexW, w2r);
    return check_launch("subcell_move_kernel");
}

// Density in sub-cell order (cell lookups staged in LDS as in force_sub_kernel,
// and the candidate positions of four stencil rows as well).
// The result goes to the reference-order row
// measures[w2r[i]] and, for a fluid row, into the sub-cell copies that the
// force pass reads: -rho in posW[i].w and p / rho^2 in velW[i].w.  Other threads read
// posW rows during this pass, but only their x, y, z; a thread writes the w of
// its own row only, after its walk.
// With masks, staged blocks also leave the in-range masks of their rows for
// the force pass (masks[r * n + i]).
template <int DIM, int HASH>
__global__ __launch_bounds__(256) void density_sub_kernel(const cmag_params P, float4* __restrict__ measures,
                                                           const uint32_t* __restrict__ w2r, float4* posW,
                                                           float4* velW, const SubCells cells, uint32_t n,
                                                           uint2* __restrict__ masks) {
    __shared__ LookupPlan lp;
    __shared__ LookupTables lt;
    const uint32_t b0 = xcd_block() * blockDim.x;
    const uint32_t i = b0 + threadIdx.x;
    __shared__ RowStage rs;
    __shared__ RowLds rows;
    const bool staged = stage_lookups<DIM>(P, posW, cells, n, b0, lp, lt);
    if (staged) stage_rows<DIM>(lp, lt, posW, rows, rs);
    if (i >= n) return;
    const float4 pi = posW[i];
    const float dens = !staged ? density_sum<CMAG_SYS_DAMBREAK, DIM, HASH>(P, pi, posW, cells)
                       : masks ? density_sum_masks<DIM>(P, pi, lp, lt, rs, &rows, posW, masks, i, n)
                               : density_sum_lookup<DIM>(P, pi, lp, lt, rs, &rows, posW);
    const float p = eos(P, dens);
    reinterpret_cast<float2*>(measures)[2 * (size_t)w2r[i]] = make_float2(dens, p);
    if (pi.w == CMAG_DB_FLUID) {
        reinterpret_cast<float*>(posW)[4 * (size_t)i + 3] = -dens;
        reinterpret_cast<float*>(velW)[4 * (size_t)i + 3] = p / (dens * dens);
    }
}

int launch_density_sub(const cmag_params& P, float* measures, const uint32_t* w2r, float* posW, float* velW,
                       const uint32_t* cellFirst, const uint32_t* pre, uint32_t n, hipStream_t s,
                       uint32_t* masks) {
    if (n == 0) return CMAG_OK;
    if (P.system != CMAG_SYS_DAMBREAK) return fail(CMAG_E_INVALID, "the sub-cell walk is DamBreak's");
    const SubCells cells{cellFirst, reinterpret_cast<const uint4*>(pre)};
    float4* m4 = reinterpret_cast<float4*>(measures);
    float4* pw = reinterpret_cast<float4*>(posW);
    float4* vw = reinterpret_cast<float4*>(velW);
    uint2* mk = reinterpret_cast<uint2*>(masks);
    dim3 g(div_up(n, 256)), b(256);
    const bool m = P.hash_order == CMAG_HASH_MORTON;
    if (P.dim == 3) {
        if (m) hipLaunchKernelGGL((density_sub_kernel<3, 1>), g, b, 0, s, P, m4, w2r, pw, vw, cells, n, mk);
        else hipLaunchKernelGGL((density_sub_kernel<3, 0>), g, b, 0, s, P, m4, w2r, pw, vw, cells, n, mk);
    } else {
        if (m) hipLaunchKernelGGL((density_sub_kernel<2, 1>), g, b, 0, s, P, m4, w2r, pw, vw, cells, n, mk);
        else hipLaunchKernelGGL((density_sub_kernel<2, 0>), g, b, 0, s, P, m4, w2r, pw, vw, cells, n, mk);
    }
    return check_launch("density_sub_kernel");
}

// Force in sub-cell order; a block whose targets share one row of cells finds
// its runs through the cell lookups staged in LDS (the same runs as the
// global walk).  CAP: neighbour queue depth; with the in-range masks only
// walked rows and blocks without staged lookups queue neighbours, and a
// shallower queue leaves more LDS to the CU (16M dam break: 4.18 vs 4.29 ms
// at depth 8 vs 16).  At depth 8 registers are held to 72 a lane so that 7
// waves share a SIMD (16M dam break: 3.61 ms, against 3.75 ms at 76 registers
// and 6 waves); at depth 16 the queues' LDS allows fewer waves anyway.
template <int DIM, int HASH, int CAP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(CAP <= 8 ? 7 : 1, 8))) void force_sub_kernel(
    const cmag_params P, float4* __restrict__ acc, const float4* __restrict__ posW, const float4* __restrict__ velW,
    const uint32_t* __restrict__ indexW, const SubCells cells, uint32_t n, const uint2* __restrict__ masks) {
    __shared__ uint32_t nbl[256 / 64][CAP][64];
    __shared__ LookupPlan lp;
    __shared__ LookupTables lt;
    const uint32_t b0 = xcd_block() * blockDim.x;
    const uint32_t i = b0 + threadIdx.x;
    const bool staged = stage_lookups<DIM>(P, posW, cells, n, b0, lp, lt);
    if (i >= n) return;
    force_body<CMAG_SYS_DAMBREAK, DIM, HASH, SubCells, CAP>(P, i, acc, nullptr, posW, velW, indexW, cells, n,
                                                            nbl[threadIdx.x >> 6], staged ? &lp : nullptr, &lt,
                                                            masks, n);
}

int launch_force_sub(const cmag_params& P, float* acc, const float* posW, const float* velW,
                     const uint32_t* indexW, const uint32_t* cellFirst, const uint32_t* pre, uint32_t n,
                     hipStream_t s, const uint32_t* masks) {
    if (n == 0) return CMAG_OK;
    if (n >= (1u << NB_SHIFT_BITS)) return fail(CMAG_E_INVALID, "force passes take at most 2^27 particles");
    if (P.system != CMAG_SYS_DAMBREAK) return fail(CMAG_E_INVALID, "the sub-cell walk is DamBreak's");
    const SubCells cells{cellFirst, reinterpret_cast<const uint4*>(pre)};
    auto run = [&](auto dim, auto hash) {
        constexpr int D = decltype(dim)::value, H = decltype(hash)::value;
        if (masks)
            hipLaunchKernelGGL((force_sub_kernel<D, H, 8>), dim3(div_up(n, 256)), dim3(256), 0, s, P,
                               reinterpret_cast<float4*>(acc), reinterpret_cast<const float4*>(posW),
                               reinterpret_cast<const float4*>(velW), indexW, cells, n,
                               reinterpret_cast<const uint2*>(masks));
        else
            hipLaunchKernelGGL((force_sub_kernel<D, H, NB_CAP>), dim3(div_up(n, 256)), dim3(256), 0, s, P,
                               reinterpret_cast<float4*>(acc), reinterpret_cast<const float4*>(posW),
                               reinterpret_cast<const float4*>(velW), indexW, cells, n, nullptr);
    };
    const bool m = P.hash_order == CMAG_HASH_MORTON;
    using I2 = std::integral_constant<int, 2>;
    using I3 = std::integral_constant<int, 3>;
    using H0 = std::integral_constant<int, 0>;
    using H1 = std::integral_constant<int, 1>;
    if (P.dim == 3) { if (m) run(I3{}, H1{}); else run(I3{}, H0{}); }
    else { if (m) run(I2{}, H1{}); else run(I2{}, H0{}); }
    return check_launch("force_sub_kernel");
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
