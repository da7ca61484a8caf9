/*
 * cmag_oracle.h - CPU oracle for the WCSPH timestep (TEST INFRASTRUCTURE).
 *
 * This library is the checker, never the product: only tests/, the smoke()
 * entry point and bench.py's cpu_baseline leg may load it.  It restates the
 * reference's CUDA loop bodies as plain C, in two precisions:
 *   *_f32  FP32 in the reference's operation order (powf, IEEE division,
 *          no FMA contraction: built with -ffp-contract=off)
 *   *_f64  the same algorithm in FP64 (accuracy anchor)
 * Hash, sort and cell ranges are integer work and exist once.
 *
 * Reference files restated (cited per function in cmag_oracle.c):
 *   DamBreak.Core/fluid_kernel.cu, fluidSystem.cpp
 *   Poiseuille.Core/poiseuilleFlowKernel.cu, poiseuilleFlowSystem.cpp
 *   Peristalsis.Core/PeristalsisKernel.cu, PeristalsisDensity.cu,
 *     PeristalsisPressureForce.cu, PeristalsisViscousForce.cu,
 *     PeristalsisIntegrate.cu, PeristalsisUtil.cu, PeristalsisSystem.cpp
 *
 * Pinning status of every function is listed at the top of cmag_oracle.c.
 */
#ifndef CMAG_ORACLE_H
#define CMAG_ORACLE_H

#include <stdint.h>
#include "../include/cmag_params.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- integer work (precision independent) ---- */
void oracle_grid_pos(const cmag_params* P, float x, float y, float z, int32_t out[3]);
uint32_t oracle_grid_hash(const cmag_params* P, const int32_t g[3]);
void oracle_calc_hash(const cmag_params* P, const float* pos4, uint32_t* hash,
                      uint32_t* index, uint32_t n);
void oracle_sort_by_key(uint32_t* keys, uint32_t* vals, uint32_t n);
void oracle_find_cell_start(uint32_t* cellStart, uint32_t* cellEnd,
                            const uint32_t* sortedHash, uint32_t n, uint32_t numCells);
/* the host harnesses need the hash of positions given in double */
void oracle_calc_hash_f64(const cmag_params* P, const double* pos4, uint32_t* hash,
                          uint32_t* index, uint32_t n);

/* ---- float stages, generated for f32 and f64 ---- */
#define CMAG_ORACLE_DECL(R, SFX)                                                         \
    void oracle_gather##SFX(R* sortedPos, R* sortedVel, const uint32_t* index,           \
                            const R* pos, const R* vel, uint32_t n);                     \
    void oracle_density##SFX(const cmag_params* P, R* measures, const R* sortedPos,      \
                             const uint32_t* cellStart, const uint32_t* cellEnd,         \
                             uint32_t n);                                                \
    void oracle_force##SFX(const cmag_params* P, R* acc, const R* measures,              \
                           const R* sortedPos, const R* sortedVel, const uint32_t* index,\
                           const uint32_t* cellStart, const uint32_t* cellEnd,           \
                           uint32_t n);                                                  \
    void oracle_pressure_force##SFX(const cmag_params* P, R* presF, const R* measures,   \
                                    const R* sortedPos, const uint32_t* index,           \
                                    const uint32_t* cellStart, const uint32_t* cellEnd,  \
                                    uint32_t n);                                         \
    void oracle_viscous_force##SFX(const cmag_params* P, R* viscF, const R* measures,    \
                                   const R* sortedPos, const R* sortedVel,               \
                                   const uint32_t* index, const uint32_t* cellStart,     \
                                   const uint32_t* cellEnd, uint32_t n);                 \
    void oracle_integrate##SFX(const cmag_params* P, R* pos, R* vel, R* velLF,           \
                               const R* acc, const R* presF, R wave_time, uint32_t n);   \
    void oracle_configure_boundary##SFX(const cmag_params* P, R* pos, R wave_height,     \
                                        uint32_t n);                                     \
    void oracle_shift_right_boundary##SFX(const cmag_params* P, R* pos, uint32_t n);     \
    void oracle_remove_right_boundary##SFX(const cmag_params* P, R* pos, uint32_t n);     \
    void oracle_density_subset##SFX(const cmag_params* P, R* measures, const R* sortedPos, \
                                    const uint32_t* cellStart, const uint32_t* cellEnd,  \
                                    const uint32_t* sel, uint32_t m);                    \
    void oracle_force_subset##SFX(const cmag_params* P, R* acc, const R* measures,       \
                                  const R* sortedPos, const R* sortedVel,                \
                                  const uint32_t* index, const uint32_t* cellStart,      \
                                  const uint32_t* cellEnd, const uint32_t* sel,          \
                                  uint32_t m);                                           \
    void oracle_density_vjp##SFX(const cmag_params* P, R* gpos, const R* grho,           \
                                 const R* sortedPos, const uint32_t* cellStart,          \
                                 const uint32_t* cellEnd, uint32_t n);                   \
    void oracle_force_vjp##SFX(const cmag_params* P, R* gpos, R* gvel, R* grho, R* gmu,  \
                               const R* gacc, const R* measures, const R* sortedPos,     \
                               const R* sortedVel, const uint32_t* cellStart,            \
                               const uint32_t* cellEnd, uint32_t n);                     \
    void oracle_integrate_vjp##SFX(const cmag_params* P, const R* pos, R* gpos, R* gvel,  \
                                   R* gvelLF, R* gacc, R* ggrav, uint32_t n);

CMAG_ORACLE_DECL(float, _f32)
CMAG_ORACLE_DECL(double, _f64)
#undef CMAG_ORACLE_DECL

/* whole update() restated; arrays are the System object's buffers.
 * acc: DamBreak/Poiseuille acceleration, Peristalsis viscous force.
 * presF: Peristalsis pressure force (unused otherwise). */
typedef struct oracle_state_f32 {
    float *pos, *vel, *velLF, *acc, *presF, *measures, *sortedPos, *sortedVel;
    uint32_t *hash, *index, *cellStart, *cellEnd;
    uint32_t n, numCells;
    float elapsed, time_shift, time_relax, wave_height;
} oracle_state_f32;
typedef struct oracle_state_f64 {
    double *pos, *vel, *velLF, *acc, *presF, *measures, *sortedPos, *sortedVel;
    uint32_t *hash, *index, *cellStart, *cellEnd;
    uint32_t n, numCells;
    double elapsed, time_shift, time_relax, wave_height;
} oracle_state_f64;

void oracle_update_f32(cmag_params* P, oracle_state_f32* S);
void oracle_update_f64(cmag_params* P, oracle_state_f64* S);

/* reference initial states (reset()) and the Peristalsis mass probe */
uint32_t oracle_dambreak_count(uint32_t fx, uint32_t fy, uint32_t fz, int32_t bo,
                               uint32_t gx, uint32_t gy, uint32_t gz);
void oracle_dambreak_init(const cmag_params* P, float* pos4, uint32_t n);
void oracle_poiseuille_init(const cmag_params* P, float* pos4, uint32_t n);
void oracle_peristalsis_init(const cmag_params* P, float* pos4, uint32_t n);
float oracle_peristalsis_mass_sum(const cmag_params* P, const float* pos4, uint32_t n);

/* threads the OpenMP loops use (0 = runtime default) */
void oracle_set_threads(int n);
int oracle_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif /* CMAG_ORACLE_H */
