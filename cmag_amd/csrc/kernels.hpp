// kernels.hpp - launchers shared between the translation units of libcmag_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>

#include "cmag_params.h"

namespace cmag {

struct StepState;

int launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);
// the demos' colour buffer of a system's kind (render.hip)
int launch_fill_colors(const cmag_params& P, const float* pos4, float* color4, uint32_t n, hipStream_t s);
int launch_copy_u32(uint32_t* dst, const uint32_t* src, size_t n, hipStream_t s);
int launch_calc_hash(const cmag_params& P, const float* pos4, uint32_t* hash, uint32_t* index,
                     uint32_t n, hipStream_t s);
// with bucket (and P): also the sub-cell bucket byte of every sorted row
int launch_reorder(uint32_t* cellStart, uint32_t* cellEnd, float* sortedPos, float* sortedVel,
                   const uint32_t* hash, const uint32_t* index, const float* oldPos,
                   const float* oldVel, uint32_t n, uint32_t numCells, hipStream_t s,
                   const cmag_params* P = nullptr, uint8_t* bucket = nullptr);

// out = exclusive_scan(in) (in place allowed); ws holds scan_levels_words(len) words
size_t scan_levels_words(size_t len);
int exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t len, uint32_t* ws, hipStream_t s);
// out = exclusive_scan of the cell counts (cellEnd - cellStart, 0 for an
// empty cell) without materialising them; ws as above
int exclusive_scan_cell_counts(const uint32_t* cellStart, const uint32_t* cellEnd, uint32_t* first,
                               size_t numCells, uint32_t* ws, hipStream_t s);

// first[c] = number of particles in cells < c, first[C] = n (from the cell
// ranges); ws holds scan_levels_words(C) words
int launch_cell_offsets(const uint32_t* cellStart, const uint32_t* cellEnd, uint32_t* first,
                        uint32_t numCells, uint32_t n, uint32_t* ws, hipStream_t s);

// stable LSD radix sort.  Sorts (keys, vals) using (alt_keys, alt_vals) as the
// ping-pong buffers; *result_in_alt tells where the sorted pair ended up.
size_t radix_sort_workspace_bytes(uint32_t n);
int radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* alt_keys, uint32_t* alt_vals,
                     uint32_t n, int key_bits, void* workspace, size_t workspace_bytes,
                     bool* result_in_alt, hipStream_t s);

// neighbour passes.  cellFirst (optional, from launch_cell_offsets) selects the
// compact-offset walker; without it the reference cellStart/cellEnd are walked.
int launch_density(const cmag_params& P, float* measures, const float* sortedPos,
                   const uint32_t* cellStart, const uint32_t* cellEnd, uint32_t n, hipStream_t s,
                   const uint32_t* cellFirst = nullptr);
// sub-cell walk order (DamBreak): build the internal order from the reference
// one (pre: 4 uint32 per cell, cleared here; bucket: the rows' buckets from
// launch_reorder, padded to whole 16-byte words + 16), then density and force
// over it; density writes the reference-order measures through w2r and leaves
// -rho (fluid rows) in posW.w and p / rho^2 in velW.w for the force pass
int launch_subcell_build(const cmag_params& P, const uint32_t* hash, const uint32_t* index, const float* sortedPos,
                         const float* sortedVel, const uint32_t* cellFirst, uint32_t n, uint32_t numCells,
                         uint32_t* pre, const uint8_t* bucket, float* posW, float* velW, uint32_t* indexW,
                         uint32_t* w2r, hipStream_t s);
// masks (optional, 2 * MASK_ROWS words per particle): density leaves the
// in-range masks of staged blocks there and force reads them
int launch_density_sub(const cmag_params& P, float* measures, const uint32_t* w2r, float* posW, float* velW,
                       const uint32_t* cellFirst, const uint32_t* pre, uint32_t n, hipStream_t s,
                       uint32_t* masks = nullptr);
int launch_force_sub(const cmag_params& P, float* acc, const float* posW, const float* velW,
                     const uint32_t* indexW, const uint32_t* cellFirst, const uint32_t* pre, uint32_t n,
                     hipStream_t s, const uint32_t* masks = nullptr);
int launch_force(const cmag_params& P, float* acc, const float* measures, const float* sortedPos,
                 const float* sortedVel, const uint32_t* index, const uint32_t* cellStart,
                 const uint32_t* cellEnd, uint32_t n, hipStream_t s,
                 const uint32_t* cellFirst = nullptr);
int launch_peristalsis_forces(const cmag_params& P, float* viscF, float* presF, const float* measures,
                              const float* sortedPos, const float* sortedVel, const uint32_t* index,
                              const uint32_t* cellStart, const uint32_t* cellEnd, uint32_t n,
                              hipStream_t s, const uint32_t* cellFirst = nullptr);
int launch_pressure_force(const cmag_params& P, float* presF, const float* measures,
                          const float* sortedPos, const uint32_t* index, const uint32_t* cellStart,
                          const uint32_t* cellEnd, uint32_t n, hipStream_t s);
int launch_viscous_force(const cmag_params& P, float* viscF, const float* measures,
                         const float* sortedPos, const float* sortedVel, const uint32_t* index,
                         const uint32_t* cellStart, const uint32_t* cellEnd, uint32_t n,
                         hipStream_t s);
// with a state record (step_state.hpp) the wave time and configuration phase
// come from it on the device, else from wave_time and P
int launch_integrate(const cmag_params& P, float* pos, float* vel, float* velLF, const float* acc,
                     const float* presF, float wave_time, uint32_t n, hipStream_t s,
                     const StepState* state = nullptr);
int launch_configure_boundary(const cmag_params& P, float* pos, float wave_height, uint32_t n,
                              hipStream_t s);
// advance one system's state record by one step (one thread), then move the
// walls to the record's configure_height when it says so
int launch_step_prologue(const cmag_params& P, StepState* state, hipStream_t s);
int launch_configure_boundary_from_state(const cmag_params& P, float* pos, const StepState* state, uint32_t n,
                                         hipStream_t s);
int launch_shift_right_boundary(const cmag_params& P, float* pos, uint32_t n, hipStream_t s);
int launch_remove_right_boundary(const cmag_params& P, float* pos, uint32_t n, hipStream_t s);

// adjoints of the Poiseuille step (adjoint.hip)
int launch_integrate_vjp(const cmag_params& P, const float* pos, float* gpos, float* gvel, float* gvelLF,
                         float* gacc, double* ggrav, uint32_t n, hipStream_t s);
int launch_gather_f4(float* dst, const float* src, const uint32_t* index, uint32_t n, hipStream_t s);
int launch_force_vjp(const cmag_params& P, float* gposS, float* gvelS, float* grho, float* gmu, double* gmu_acc,
                     const float* gaccS, const float* measures, const float* sortedPos, const float* sortedVel,
                     const uint32_t* cellStart, const uint32_t* cellEnd, const uint32_t* cellFirst, uint32_t n,
                     hipStream_t s);
int launch_density_vjp(const cmag_params& P, float* gposS, const float* grho, const float* sortedPos,
                       const uint32_t* cellStart, const uint32_t* cellEnd, const uint32_t* cellFirst, uint32_t n,
                       const uint32_t* index, const float* gvelS, float* gpos, float* gvelLF, hipStream_t s);

}  // namespace cmag
