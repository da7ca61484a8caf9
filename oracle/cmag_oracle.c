/*
 * cmag_oracle.c - CPU oracle of the WCSPH timestep.  TEST INFRASTRUCTURE:
 * loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg, and only as the checker (never as the measured or shipped path).
 *
 * How each function is pinned (checks live in tests/test_oracle.py):
 *   oracle_grid_pos / grid_hash / calc_hash
 *       closed form: particles placed at known integer cells (interior,
 *       exact cell-boundary multiples, negative cells that wrap, last cell),
 *       both rounding rules; Morton order checked as a bijection whose
 *       restriction to each axis is the textbook bit spread.
 *   oracle_sort_by_key
 *       numpy.argsort(kind="stable") (library routine) on random, duplicate-
 *       heavy, already-sorted, reversed and empty inputs.
 *   oracle_find_cell_start / gather
 *       numpy.unique(return_index, return_counts) partition of [0,N);
 *       gather equals fancy indexing.
 *   oracle_density_*
 *       reference-held known answer: the Poiseuille mass constant
 *       1000/3381320880.551724 (poiseuilleFlowSystem.cpp:49) is rho0 over the
 *       lattice kernel sum, so interior particles of the reference lattice
 *       must give rho = 1000 (to FP32 accumulation); isolated particle gives
 *       m*W(0); 3-D normalisation integrates to 1 over a fine lattice;
 *       O(N^2) brute force (no cell list) on random inputs; periodic image
 *       invariance (translating every particle by one period).
 *   oracle_force_* / pressure_force_* / viscous_force_*
 *       momentum conservation (sum m*a = 0 for a periodic fluid-only box, which
 *       catches a wrong sign or a dropped symmetric term); zero force on a
 *       perfect lattice at uniform state; SPH consistency: a linear pressure
 *       field gives a = -grad(p)/rho and a quadratic shear flow
 *       u = y^2 gives the Morris viscous acceleration 2*nu; Lennard-Jones term
 *       against a single wall particle in closed form; O(N^2) brute force.
 *       DamBreak (tests/test_oracle_pins.py): the linear-pressure consistency
 *       check on 3-D and 2-D lattices (h = 2dx and 1.3dx) fixes the gradient
 *       normalisations 21/(8 pi h^4) and 7/(2 pi h^3); two-particle closed
 *       forms with exact inputs fix the Monaghan term (nu = 2 alpha h c /
 *       (rho_i + rho_j), the 0.001 h^2 regulariser, zero for separating
 *       pairs); the viscosity conserves momentum and dissipates energy.
 *       Every such pin rejects a planted mistake (ORACLE_MUTATE builds).
 *   Tait equation of state
 *       closed forms p(rho0) = 0, p(2^(1/7) rho0) = B, p(2 rho0) = 127 B and
 *       dp/drho(rho0) = B gamma / rho0, densities set through an isolated
 *       particle's own kernel value; planted exponent and B mistakes fail.
 *   oracle_integrate_* / configure_boundary / shift / remove
 *       closed-form single-particle updates with dyadic numbers (exact in
 *       binary), wrap and clamp cases, the analytic wall sinusoid.
 *   oracle_update_*
 *       end-to-end: Poiseuille transient series (Poiseuille.Report/
 *       XVelocityYPosition.m:12) within 1% L2 on the channel profile.
 *   init functions
 *       particle counts against the reference's numParticles formulas and
 *       lattice spacing / wall-layer geometry.
 * Nothing in the reference itself pins numbers (it has no tests or golden
 * files); the reference cannot be compiled here (CUDA only), so no
 * reference build backs these checks.
 */
#include "cmag_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)

/* ------------------------------------------------------------------------ */
/* calcGridPos: the two rounding rules, each exactly as the reference spells it
 *   DamBreak    floor((p - o) / cellSize)          fluid_kernel.cu:19-25
 *   Poiseuille  floor((p - o) * 0.5f / radius)     poiseuilleFlowKernel.cu:18-24
 *   Peristalsis same as Poiseuille                 PeristalsisKernel.cu:18-24 */
void oracle_grid_pos(const cmag_params* P, float x, float y, float z, int32_t out[3]) {
    float p[3] = {x, y, z};
    for (int a = 0; a < 3; ++a) {
        float v;
        if (P->gridpos_rule == CMAG_GRIDPOS_DIV_CELL)
            v = (p[a] - P->world_origin[a]) / P->cell_size[a];
        else
            v = (p[a] - P->world_origin[a]) * 0.5f / P->particle_radius;
        out[a] = (int32_t)floorf(v);
    }
}

/* calcGridHash, fluid_kernel.cu:27-32: mask each axis with gridSize-1 (the
 * two's-complement wrap), then linear order.  Z order: the low xbits bits of
 * x, followed by the interleave of (x >> xbits, y, z) bit by bit, x first,
 * skipping axes that ran out of bits -- a Morton curve over bricks of
 * 2^xbits x 1 x 1 cells (xbits = 0 is the plain curve). */
uint32_t oracle_grid_hash(const cmag_params* P, const int32_t g[3]) {
    uint32_t c[3];
    for (int a = 0; a < 3; ++a) c[a] = (uint32_t)g[a] & (P->grid_size[a] - 1u);
    if (P->hash_order == CMAG_HASH_LINEAR)
        return c[2] * P->grid_size[1] * P->grid_size[0] + c[1] * P->grid_size[0] + c[0];
    uint32_t xb = P->morton_xbits < P->grid_bits[0] ? P->morton_xbits : P->grid_bits[0];
    uint32_t coord[3] = {c[0] >> xb, c[1], c[2]};
    uint32_t bits[3] = {P->grid_bits[0] - xb, P->grid_bits[1], P->grid_bits[2]};
    uint32_t key = c[0] & ((1u << xb) - 1u), bit = xb;
    uint32_t maxb = bits[0];
    if (bits[1] > maxb) maxb = bits[1];
    if (bits[2] > maxb) maxb = bits[2];
    for (uint32_t b = 0; b < maxb; ++b)
        for (int a = 0; a < 3; ++a)
            if (b < bits[a]) key |= ((coord[a] >> b) & 1u) << bit++;
    return key;
}

/* calcHashD, fluid_kernel.cu:34-48 */
void oracle_calc_hash(const cmag_params* P, const float* pos4, uint32_t* hash,
                      uint32_t* index, uint32_t n) {
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        int32_t g[3];
        oracle_grid_pos(P, pos4[4 * i], pos4[4 * i + 1], pos4[4 * i + 2], g);
        hash[i] = oracle_grid_hash(P, g);
        index[i] = (uint32_t)i;
    }
}

void oracle_calc_hash_f64(const cmag_params* P, const double* pos4, uint32_t* hash,
                          uint32_t* index, uint32_t n) {
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        int32_t g[3];
        oracle_grid_pos(P, (float)pos4[4 * i], (float)pos4[4 * i + 1], (float)pos4[4 * i + 2], g);
        hash[i] = oracle_grid_hash(P, g);
        index[i] = (uint32_t)i;
    }
}

/* thrust::sort_by_key on (hash, index) is a stable ascending sort
 * (fluidSystem.cu:94-99); restated as an LSD radix sort, 8-bit digits. */
void oracle_sort_by_key(uint32_t* keys, uint32_t* vals, uint32_t n) {
    if (n < 2) return;
    uint32_t maxk = 0;
    for (uint32_t i = 0; i < n; ++i) if (keys[i] > maxk) maxk = keys[i];
    uint32_t* k2 = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t* v2 = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t *ka = keys, *va = vals, *kb = k2, *vb = v2;
    for (uint32_t shift = 0; shift < 32 && (shift == 0 || (maxk >> shift) != 0); shift += 8) {
        size_t cnt[257];
        memset(cnt, 0, sizeof(cnt));
        for (uint32_t i = 0; i < n; ++i) cnt[((ka[i] >> shift) & 255u) + 1]++;
        for (int d = 0; d < 256; ++d) cnt[d + 1] += cnt[d];
        for (uint32_t i = 0; i < n; ++i) {
            size_t dst = cnt[(ka[i] >> shift) & 255u]++;
            kb[dst] = ka[i];
            vb[dst] = va[i];
        }
        uint32_t* t = ka; ka = kb; kb = t;
        t = va; va = vb; vb = t;
    }
    if (ka != keys) {
        memcpy(keys, ka, sizeof(uint32_t) * n);
        memcpy(vals, va, sizeof(uint32_t) * n);
    }
    free(k2);
    free(v2);
}

/* reorderDataAndFindCellStartD's run detection, fluid_kernel.cu:50-89, plus
 * the cudaMemset(cellStart, 0xff) of its launcher (fluidSystem.cu:164).
 * cellEnd of empty cells is left untouched, as in the reference. */
void oracle_find_cell_start(uint32_t* cellStart, uint32_t* cellEnd,
                            const uint32_t* sortedHash, uint32_t n, uint32_t numCells) {
    for (uint32_t c = 0; c < numCells; ++c) cellStart[c] = CMAG_CELL_EMPTY;
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t h = sortedHash[i];
        if (i == 0 || h != sortedHash[i - 1]) {
            cellStart[h] = i;
            if (i > 0) cellEnd[sortedHash[i - 1]] = i;
        }
        if (i == n - 1) cellEnd[h] = i + 1;
    }
}

/* ---- FP32 reference-order stages ---- */
#define R float
#define SFX _f32
#define SQRT sqrtf
#define POW powf
#define SIN sinf
#define COS cosf
#define FMIN fminf
#define PI_R 3.141592654f
#include "oracle_stages.inc"
#undef R
#undef SFX
#undef SQRT
#undef POW
#undef SIN
#undef COS
#undef FMIN
#undef PI_R

/* ---- FP64 stages ---- */
#define R double
#define SFX _f64
#define SQRT sqrt
#define POW pow
#define SIN sin
#define COS cos
#define FMIN fmin
#define PI_R 3.14159265358979323846
#include "oracle_stages.inc"
#undef R
#undef SFX
#undef SQRT
#undef POW
#undef SIN
#undef COS
#undef FMIN
#undef PI_R

/* ------------------------------------------------------------------------ */
/* update(): the stage order of
 *   DamBreakSystem::update        fluidSystem.cpp:257-319
 *   PoiseuilleFlowSystem::update  poiseuilleFlowSystem.cpp:212-270
 *   PeristalsisSystem::Update     PeristalsisSystem.cpp:320-407 (with its
 *                                 boundary-configuration state machine :330-343) */
#define ORACLE_UPDATE(R, SFX, HASHFN)                                                       \
    void oracle_update##SFX(cmag_params* P, oracle_state##SFX* S) {                       \
        R t_w = 0;                                                                          \
        if (P->system == CMAG_SYS_PERISTALSIS) {                                            \
            if (P->is_boundary_configuration) {                                             \
                S->time_shift += (R)P->delta_time;                                          \
                if (S->wave_height < (R)P->amplitude) {                                     \
                    oracle_configure_boundary##SFX(P, S->pos, S->wave_height, S->n);        \
                    S->wave_height += (R)P->delta_time * (R)powf(10.0f, -3.0f);            \
                } else {                                                                    \
                    S->time_relax -= (R)P->delta_time;                                      \
                    if (S->time_relax < 0) P->is_boundary_configuration = 0;                \
                }                                                                           \
            }                                                                               \
            t_w = P->is_boundary_configuration ? (R)0 : S->elapsed - S->time_shift;         \
        }                                                                                   \
        HASHFN(P, S->pos, S->hash, S->index, S->n);                                         \
        oracle_sort_by_key(S->hash, S->index, S->n);                                        \
        oracle_find_cell_start(S->cellStart, S->cellEnd, S->hash, S->n, S->numCells);       \
        oracle_gather##SFX(S->sortedPos, S->sortedVel, S->index, S->pos, S->velLF, S->n);   \
        oracle_density##SFX(P, S->measures, S->sortedPos, S->cellStart, S->cellEnd, S->n);  \
        if (P->system == CMAG_SYS_PERISTALSIS) {                                            \
            oracle_viscous_force##SFX(P, S->acc, S->measures, S->sortedPos, S->sortedVel,   \
                                      S->index, S->cellStart, S->cellEnd, S->n);            \
            oracle_pressure_force##SFX(P, S->presF, S->measures, S->sortedPos, S->index,    \
                                       S->cellStart, S->cellEnd, S->n);                     \
        } else {                                                                            \
            oracle_force##SFX(P, S->acc, S->measures, S->sortedPos, S->sortedVel, S->index, \
                              S->cellStart, S->cellEnd, S->n);                              \
        }                                                                                   \
        oracle_integrate##SFX(P, S->pos, S->vel, S->velLF, S->acc, S->presF, t_w, S->n);    \
        S->elapsed += (R)P->delta_time;                                                     \
    }

static void hash_f32(const cmag_params* P, float* pos, uint32_t* h, uint32_t* idx, uint32_t n) {
    oracle_calc_hash(P, pos, h, idx, n);
}
static void hash_f64(const cmag_params* P, double* pos, uint32_t* h, uint32_t* idx, uint32_t n) {
    oracle_calc_hash_f64(P, pos, h, idx, n);
}
ORACLE_UPDATE(float, _f32, hash_f32)
ORACLE_UPDATE(double, _f64, hash_f64)

/* ------------------------------------------------------------------------ */
/* initial states */

/* DamBreakSystem ctor count, fluidSystem.cpp:34-37 */
uint32_t oracle_dambreak_count(uint32_t fx, uint32_t fy, uint32_t fz, int32_t bo,
                               uint32_t gx, uint32_t gy, uint32_t gz) {
    (void)gz;
    return fx * fy * fz + gx * (uint32_t)bo + 2u * (gy - (uint32_t)bo) * (uint32_t)bo;
}

static void put4(float* a, uint32_t i, float x, float y, float z, float w) {
    a[4 * (size_t)i] = x; a[4 * (size_t)i + 1] = y; a[4 * (size_t)i + 2] = z; a[4 * (size_t)i + 3] = w;
}

/* DamBreakSystem::initFluid + initBoundaryParticles, fluidSystem.cpp:382-523 */
void oracle_dambreak_init(const cmag_params* P, float* pos4, uint32_t n) {
    const float r = P->particle_radius, spacing = r * 2.0f;
    const int bo = P->boundary_offset;
    const float hx = P->grid_size[0] * r, hy = P->grid_size[1] * r, hz = P->grid_size[2] * r;
    const float ox = P->world_origin[0], oy = P->world_origin[1], oz = P->world_origin[2];
    const uint32_t fx = P->fluid_size[0], fy = P->fluid_size[1], fz = P->fluid_size[2];
    memset(pos4, 0, sizeof(float) * 4 * (size_t)n);
    for (uint32_t z = 0; z < fz; z++)
        for (uint32_t y = 0; y < fy; y++)
            for (uint32_t x = 0; x < fx; x++) {
                uint32_t i = (z * fy * fx) + y * fx + x;
                if (i < n)
                    put4(pos4, i, (spacing * x) + r - hx + bo * 2 * r,
                         (spacing * y) + r - hy + bo * 2 * r, (spacing * z) + r - hz, CMAG_DB_FLUID);
            }
    uint32_t base = fx * fy * fz;
    uint32_t sx, sy;
    /* bottom type 1 */
    sx = P->grid_size[0] - (bo - 1); sy = 1;
    for (uint32_t y = 0; y < sy; y++)
        for (uint32_t x = 0; x < sx; x++)
            put4(pos4, base + y * sx + x, (spacing * x) + r + ox + (bo - 1) * 2 * r,
                 (spacing * y) + r + oy + (bo - 1) * 2 * r, r + oz, CMAG_DB_FIRST);
    base += sx * sy;
    /* left type 1 */
    sx = 1; sy = P->grid_size[1] - bo;
    for (uint32_t y = 0; y < sy; y++)
        for (uint32_t x = 0; x < sx; x++)
            put4(pos4, base + y * sx + x, (spacing * x) + r + ox + (bo - 1) * 2 * r,
                 (spacing * y) + r + oy + (bo + 0) * 2 * r, r + oz, CMAG_DB_FIRST);
    base += sx * sy;
    /* right type 1 */
    sx = 1; sy = P->grid_size[1] - bo;
    for (uint32_t y = 0; y < sy; y++)
        for (uint32_t x = 0; x < sx; x++)
            put4(pos4, base + y * sx + x, (spacing * x) + r + ox + (fx + bo) * 2 * r,
                 (spacing * y) + r + oy + bo * 2 * r, r + oz, CMAG_DB_RIGHT_FIRST);
    base += sx * sy;
    /* bottom type 2 */
    sx = P->grid_size[0]; sy = bo - 1;
    for (uint32_t y = 0; y < sy; y++)
        for (uint32_t x = 0; x < sx; x++)
            put4(pos4, base + y * sx + x, (spacing * x) + r + ox, (spacing * y) + r + oy, r + oz,
                 CMAG_DB_SECOND);
    base += sx * sy;
    /* left type 2 */
    sx = bo - 1; sy = P->grid_size[1] - (bo - 1);
    for (uint32_t y = 0; y < sy; y++)
        for (uint32_t x = 0; x < sx; x++)
            put4(pos4, base + y * sx + x, (spacing * x) + r + ox,
                 (spacing * y) + r + oy + (bo - 1) * 2 * r, r + oz, CMAG_DB_SECOND);
    base += sx * sy;
    /* right type 2 */
    sx = bo - 1; sy = P->grid_size[1] - bo;
    for (uint32_t y = 0; y < sy; y++)
        for (uint32_t x = 0; x < sx; x++)
            put4(pos4, base + y * sx + x, (spacing * x) + r + ox + (fx + bo + 1) * 2 * r,
                 (spacing * y) + r + oy + (bo) * 2 * r, r + oz, CMAG_DB_RIGHT_SECOND);
}

/* PoiseuilleFlowSystem::initFluid + initBoundaryParticles,
 * poiseuilleFlowSystem.cpp:325-383 */
void oracle_poiseuille_init(const cmag_params* P, float* pos4, uint32_t n) {
    const float r = P->particle_radius, spacing = r * 2.0f;
    const int bo = P->boundary_offset;
    const float hx = P->grid_size[0] * r, hy = P->grid_size[1] * r, hz = P->grid_size[2] * r;
    const float ox = P->world_origin[0], oy = P->world_origin[1], oz = P->world_origin[2];
    const uint32_t fx = P->fluid_size[0], fy = P->fluid_size[1], fz = P->fluid_size[2];
    memset(pos4, 0, sizeof(float) * 4 * (size_t)n);
    for (uint32_t z = 0; z < fz; z++)
        for (uint32_t y = 0; y < fy; y++)
            for (uint32_t x = 0; x < fx; x++) {
                uint32_t i = (z * fy * fx) + y * fx + x;
                if (i < n)
                    put4(pos4, i, (spacing * x) + r - hx, (spacing * y) + r - hy + bo * 2 * r,
                         (spacing * z) + r - hz, CMAG_PO_FLUID);
            }
    uint32_t base = fx * fy * fz;
    uint32_t sx = P->grid_size[0], sy = (uint32_t)bo;
    for (uint32_t y = 0; y < sy; y++)
        for (uint32_t x = 0; x < sx; x++)
            put4(pos4, base + y * sx + x, (spacing * x) + r + ox, (spacing * y) + r + oy, r + oz,
                 CMAG_PO_WALL);
    base += sx * sy;
    for (uint32_t y = 0; y < sy; y++)
        for (uint32_t x = 0; x < sx; x++)
            put4(pos4, base + y * sx + x, (spacing * x) + r + ox,
                 (spacing * y) + r + oy + bo * 2 * r + fy * 2.0f * r, r + oz, CMAG_PO_WALL);
}

/* PeristalsisSystem::initFluid + initBoundaryParticles, PeristalsisSystem.cpp:255-317 */
void oracle_peristalsis_init(const cmag_params* P, float* pos4, uint32_t n) {
    const float r = P->particle_radius, spacing = r * 2.0f;
    const int bo = P->boundary_offset;
    const float hx = P->grid_size[0] * r, hz = P->grid_size[2] * r;
    const float ox = P->world_origin[0], oz = P->world_origin[2];
    const uint32_t fx = P->fluid_size[0], fy = P->fluid_size[1], fz = P->fluid_size[2];
    memset(pos4, 0, sizeof(float) * 4 * (size_t)n);
    for (int z = 0; z < (int)fz; z++)
        for (int y = 0; y < (int)fy; y++)
            for (int x = 0; x < (int)fx; x++) {
                uint32_t i = (z * fy * fx) + y * fx + x;
                if (i < n)
                    put4(pos4, i, (spacing * x) + r - hx, -1.0f * fy * r + (spacing * y) + r,
                         (spacing * z) + r - hz, CMAG_PE_FLUID);
            }
    uint32_t base = fx * fy * fz;
    uint32_t sx = fx, sy = (uint32_t)bo;
    for (uint32_t y = 0; y < sy; y++)
        for (uint32_t x = 0; x < sx; x++) {
            float j = r * (2 * x + 1);
            put4(pos4, base + y * sx + x, j + ox, -1.0f * fy * r - (spacing * y) - r, r + oz,
                 2.0f * (1.0f + y));
        }
    base += sx * sy;
    for (uint32_t y = 0; y < sy; y++)
        for (uint32_t x = 0; x < sx; x++) {
            float j = r * (2 * x + 1);
            put4(pos4, base + y * sx + x, j + ox, fy * r + (spacing * y) + r, r + oz,
                 -2.0f * (1.0f + y));
        }
}

/* PeristalsisSystem::CalculateMass, PeristalsisSystem.cpp:238-253: kernel sum at
 * the probe particle gridSize.x/2 over every particle, x-y distance, no images */
float oracle_peristalsis_mass_sum(const cmag_params* P, const float* pos4, uint32_t n) {
    float x = pos4[(P->grid_size[0] / 2) * 4 + 0];
    float y = pos4[(P->grid_size[0] / 2) * 4 + 1];
    float sum = 0.0f;
    for (uint32_t i = 0; i < n; i++) {
        float rx = x - pos4[4 * (size_t)i], ry = y - pos4[4 * (size_t)i + 1];
        float dist = sqrtf(rx * rx + ry * ry);
        float q = dist / P->smoothing_radius;
        float coeff = 7.0f / 4 / 3.141592654f / powf(P->smoothing_radius, 2);
        if (q < 2) sum += coeff * (powf(1 - 0.5f * q, 4) * (2 * q + 1));
    }
    return sum;
}

void oracle_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int oracle_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
