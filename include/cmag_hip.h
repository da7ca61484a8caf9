/*
 * cmag_hip.h - C ABI of libcmag_hip.so, the MI355X WCSPH timestep.
 *
 * Plain C: device pointers, counts, a parameter block (cmag_params.h) and a
 * stream handle; every entry point returns an int status (CMAG_OK = 0).
 * Launchers are asynchronous on the given stream (NULL = the default stream)
 * and never synchronise; host-array calls (set/get array) synchronise the
 * object's stream.  Errors never exit the process: cmag_last_error() returns
 * the message of the last failure on the calling thread.
 *
 * Each launcher replaces one extern "C" launcher of the reference's
 * System*.cu files (cited per entry).  The reference reused the same symbol
 * names in all three libraries (sortParticles, setParameters, ...) and kept
 * parameters in a process-global __constant__; here names carry a cmag_
 * prefix, the system is selected by cmag_params.system, and the parameter
 * block travels with every call, so instances and systems can coexist.
 */
#ifndef CMAG_HIP_H
#define CMAG_HIP_H

#include <stddef.h>
#include <stdint.h>
#include "cmag_params.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef void* cmag_stream_t;             /* a hipStream_t */
typedef struct cmag_system cmag_system;  /* opaque System object */

enum cmag_status {
    CMAG_OK = 0,
    CMAG_E_INVALID = -1, /* bad argument (null pointer, size, grid not a power of two) */
    CMAG_E_HIP = -2,     /* a HIP runtime call failed */
    CMAG_E_SIZE = -3,    /* scratch or workspace too small */
    CMAG_E_STATE = -4    /* call not valid in the object's current state */
};

/* particle arrays of a System object (float4 per particle) */
enum cmag_array {
    CMAG_ARRAY_POSITION = 0,
    CMAG_ARRAY_VELOCITY = 1,
    CMAG_ARRAY_MEASURES = 2,         /* (rho, p, ., .) in sorted order */
    CMAG_ARRAY_ACCELERATION = 3,     /* DamBreak/Poiseuille acceleration; Peristalsis viscous force */
    CMAG_ARRAY_VELOCITYLEAPFROG = 4,
    CMAG_ARRAY_PRESSUREFORCE = 5,    /* Peristalsis only */
    CMAG_ARRAY_SORTEDPOS = 6,
    CMAG_ARRAY_SORTEDVEL = 7,
    CMAG_ARRAY_PREDICTEDPOSITION = 8,/* Peristalsis only (kept for setArray parity) */
    /* uint32 arrays */
    CMAG_ARRAY_HASH = 16,
    CMAG_ARRAY_INDEX = 17,
    CMAG_ARRAY_CELLSTART = 18,
    CMAG_ARRAY_CELLEND = 19
};

/* ---- library ---------------------------------------------------------- */
const char* cmag_last_error(void);
int cmag_version(void);
/* sizeof(cmag_params) as compiled into the library (binding check) */
int cmag_params_size(void);

/* ---- parameter blocks (host only, no device work) ------------------------ */
/* DamBreakSystem ctor, DamBreak.Core/fluidSystem.cpp:17-77 */
int cmag_params_dambreak(cmag_params* out, uint32_t fx, uint32_t fy, uint32_t fz,
                         int32_t boundaryOffset, uint32_t gx, uint32_t gy, uint32_t gz,
                         float particleRadius);
/* PoiseuilleFlowSystem ctor, Poiseuille.Core/poiseuilleFlowSystem.cpp:21-67 */
int cmag_params_poiseuille(cmag_params* out, uint32_t fx, uint32_t fy, uint32_t fz,
                           int32_t boundaryOffset, uint32_t gx, uint32_t gy, uint32_t gz,
                           float particleRadius);
/* PeristalsisSystem ctor, Peristalsis.Core/PeristalsisSystem.cpp:17-80 (mass is
 * set by cmag_system_reset from the lattice, PeristalsisSystem.cpp:232) */
int cmag_params_peristalsis(cmag_params* out, float deltaTime, uint32_t fx, uint32_t fy,
                            uint32_t fz, float amplitude, float wave_speed, float soundspeed,
                            float gravx, float gravy, float gravz, int32_t boundaryOffset,
                            uint32_t gx, uint32_t gy, uint32_t gz, float radius);
/* recompute the derived fields (grid_bits, Morton masks, kernel coefficients,
 * wall planes, clamp borders) after editing a block by hand */
int cmag_params_finalize(cmag_params* p);
uint32_t cmag_params_num_cells(const cmag_params* p);
/* number of key bits of the cell hash: ceil(log2(num_cells)) */
int cmag_params_key_bits(const cmag_params* p);

/* the reference reset() lattice for p->system (host memory, no device work):
 * DamBreakSystem::initFluid/initBoundaryParticles fluidSystem.cpp:382-523,
 * PoiseuilleFlowSystem's poiseuilleFlowSystem.cpp:325-383,
 * PeristalsisSystem's PeristalsisSystem.cpp:255-317 */
int cmag_reference_initial_state(const cmag_params* p, uint32_t n, float* pos4);
/* PeristalsisSystem::CalculateMass's kernel sum, PeristalsisSystem.cpp:238-253 */
float cmag_peristalsis_lattice_sum(const cmag_params* p, const float* pos4, uint32_t n);

/* ---- stage launchers (device pointers; asynchronous) --------------------- */
/* calcHash: fluidSystem.cu:133-148 (+ poiseuilleFlowSystem.cu:110-125,
 * PeristalsisSystem.cu:99-114).  hash[i] = cell key of pos[i], index[i] = i. */
int cmag_calc_hash(const cmag_params* p, const float* pos4, uint32_t* hash,
                   uint32_t* index, uint32_t n, cmag_stream_t stream);

/* sortParticles: fluidSystem.cu:94-99 (thrust::sort_by_key).  Stable
 * ascending LSD radix sort of keys < 2^key_bits, in place, with a scratch
 * buffer of at least cmag_sort_scratch_bytes(n, key_bits) bytes. */
size_t cmag_sort_scratch_bytes(uint32_t n, int key_bits);
int cmag_sort_by_key(uint32_t* keys, uint32_t* vals, uint32_t n, int key_bits, void* scratch,
                     size_t scratch_bytes, cmag_stream_t stream);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  cmag_sort_by_key with its scratch in a file-scope device array
 * of cmag_sort_noalloc_scratch_bytes() bytes (35,651,840: alternate keys and
 * values plus the digit tables for CMAG_SORT_NOALLOC_MAX_KEYS keys), held by
 * every GPU the library runs on until the process ends.  Calls no hipMalloc,
 * hipFree or allocating library routine; n above the limit returns CMAG_E_SIZE.
 * Sorts on the current device (the stream must belong to it).  All calls on
 * one GPU share that GPU's array, so calls that overlap on one GPU (other
 * streams, other host threads) race; calls in one stream order are safe.
 * Nothing in the array needs to start at zero. */
#define CMAG_SORT_NOALLOC_MAX_KEYS 4194304u
size_t cmag_sort_noalloc_scratch_bytes(void);
int cmag_sort_by_key_noalloc(uint32_t* keys, uint32_t* vals, uint32_t n, int key_bits, cmag_stream_t stream);

/* reorderDataAndFindCellStart: fluidSystem.cu:150-188 (incl. its cudaMemset of
 * cellStart).  cellEnd of empty cells is not written. */
int cmag_reorder_find_cell_start(uint32_t* cellStart, uint32_t* cellEnd, float* sortedPos,
                                 float* sortedVel, const uint32_t* hash, const uint32_t* index,
                                 const float* oldPos, const float* oldVel, uint32_t n,
                                 uint32_t numCells, cmag_stream_t stream);

/* density + equation of state, by p->system:
 *   calculateDamBreakDensity    fluidSystem.cu:190-230
 *   calculatePoiseuilleDensity  poiseuilleFlowSystem.cu:167-204
 *   computeDensityVariation     PeristalsisSystem.cu:163-199 */
int cmag_density(const cmag_params* p, float* measures, const float* sortedPos,
                 const uint32_t* cellStart, const uint32_t* cellEnd, uint32_t n,
                 cmag_stream_t stream);

/* in-range neighbours (j != i, |r| < 2h) of every sorted particle, walked
 * exactly as the density pass walks them.  A measurement aid with no
 * reference counterpart: it gives the neighbour-pair count per pass. */
int cmag_count_neighbours(const cmag_params* p, uint32_t* counts, const float* sortedPos,
                          const uint32_t* cellStart, const uint32_t* cellEnd, uint32_t n,
                          cmag_stream_t stream);

/* fused force pass scattered to acc[index[i]]:
 *   calcAndApplyAcceleration         fluidSystem.cu:232-272
 *   calculatePoiseuilleAcceleration  poiseuilleFlowSystem.cu:206-247 */
int cmag_force(const cmag_params* p, float* acc, const float* measures, const float* sortedPos,
               const float* sortedVel, const uint32_t* index, const uint32_t* cellStart,
               const uint32_t* cellEnd, uint32_t n, cmag_stream_t stream);

/* computePressureForce, PeristalsisSystem.cu:245-283 */
int cmag_pressure_force(const cmag_params* p, float* presF, const float* measures,
                        const float* sortedPos, const uint32_t* index,
                        const uint32_t* cellStart, const uint32_t* cellEnd, uint32_t n,
                        cmag_stream_t stream);

/* computeViscousForce, PeristalsisSystem.cu:201-243 */
int cmag_viscous_force(const cmag_params* p, float* viscF, const float* measures,
                       const float* sortedPos, const float* sortedVel, const uint32_t* index,
                       const uint32_t* cellStart, const uint32_t* cellEnd, uint32_t n,
                       cmag_stream_t stream);

/* integrate, by p->system (presF and wave_time are Peristalsis only):
 *   integrateSystem            fluidSystem.cu:114-131
 *   integratePoiseuilleSystem  poiseuilleFlowSystem.cu:84-101
 *   computeCoordinates         PeristalsisSystem.cu:308-329 */
int cmag_integrate(const cmag_params* p, float* pos, float* vel, float* velLF, const float* acc,
                   const float* presF, float wave_time, uint32_t n, cmag_stream_t stream);

/* ExtConfigureBoundary, PeristalsisSystem.cu:84-97 */
int cmag_configure_boundary(const cmag_params* p, float* pos, float wave_height, uint32_t n,
                            cmag_stream_t stream);
/* ExtChangeRightBoundary, fluidSystem.cu:81-92 */
int cmag_shift_right_boundary(const cmag_params* p, float* pos, uint32_t n, cmag_stream_t stream);
/* ExtRemoveRightBoundary, fluidSystem.cu:101-112 */
int cmag_remove_right_boundary(const cmag_params* p, float* pos, uint32_t n, cmag_stream_t stream);

/* ---- System objects: the reference classes behind a C handle ------------- */
/* DamBreakSystem(fluidParticlesSize, boundaryOffset, gridSize, particleRadius),
 * DamBreak.Core/fluidSystem.h:9-14 */
int cmag_dambreak_create(uint32_t fx, uint32_t fy, uint32_t fz, int32_t boundaryOffset,
                         uint32_t gx, uint32_t gy, uint32_t gz, float particleRadius,
                         cmag_system** out);
/* PoiseuilleFlowSystem(...), Poiseuille.Core/poiseuilleFlowSystem.h:9-14 */
int cmag_poiseuille_create(uint32_t fx, uint32_t fy, uint32_t fz, int32_t boundaryOffset,
                           uint32_t gx, uint32_t gy, uint32_t gz, float particleRadius,
                           cmag_system** out);
/* PeristalsisSystem(...), Peristalsis.Core/PeristalsisSystem.h:9-19 */
int cmag_peristalsis_create(float deltaTime, uint32_t fx, uint32_t fy, uint32_t fz,
                            float amplitude, float wave_speed, float soundspeed, float gravx,
                            float gravy, float gravz, int32_t boundaryOffset, uint32_t gx,
                            uint32_t gy, uint32_t gz, float radius, cmag_system** out);
/* a System on an arbitrary parameter block with n particles; the caller fills
 * the arrays with cmag_system_set_array (reset() is not available) */
int cmag_system_create(const cmag_params* p, uint32_t n, cmag_system** out);
int cmag_system_destroy(cmag_system* s);

/* reset(): fluidSystem.cpp:361-380, poiseuilleFlowSystem.cpp:311-323,
 * PeristalsisSystem.cpp:213-237 (Reset) */
int cmag_system_reset(cmag_system* s);
/* update(): fluidSystem.cpp:257-319, poiseuilleFlowSystem.cpp:212-270,
 * PeristalsisSystem.cpp:320-407 (Update) */
int cmag_system_update(cmag_system* s);
/* update(dt): sets the time step, then update() */
int cmag_system_update_dt(cmag_system* s, float dt);
/* steps consecutive updates, replaying one captured HIP graph per step when
 * use_graph is non-zero (all three systems: the Peristalsis boundary-
 * configuration state machine runs on the device) */
int cmag_system_run(cmag_system* s, uint32_t steps, int use_graph);

/* setArray(array, data, start, count), fluidSystem.cpp:321-355: count
 * particles of 4 floats from host memory, starting at particle `start` */
int cmag_system_set_array(cmag_system* s, int array, const float* data, int start, int count);
/* getArray: the device-to-host counterpart (not in the reference, whose
 * harnesses copied from the raw device-pointer getters) */
int cmag_system_get_array(cmag_system* s, int array, float* data, int start, int count);
int cmag_system_get_uint_array(cmag_system* s, int array, uint32_t* data, uint32_t start,
                               uint32_t count);
/* the raw device pointer getters, fluidSystem.h:40-47 (borrowed) */
void* cmag_system_device_ptr(cmag_system* s, int array);

uint32_t cmag_system_num_particles(const cmag_system* s);
uint32_t cmag_system_num_cells(const cmag_system* s);
float cmag_system_elapsed_time(const cmag_system* s);
int cmag_system_get_params(const cmag_system* s, cmag_params* out);
/* replaces the block (the setParameters of the reference); a new block may
 * not change the particle count or the grid */
int cmag_system_set_params(cmag_system* s, const cmag_params* p);
int cmag_system_set_stream(cmag_system* s, cmag_stream_t stream);
/* Neighbour-walk order of a DamBreak system (no counterpart in the reference,
 * whose walk is the only order): with on != 0 the density and force passes
 * run over an internal copy of the sorted rows ordered inside each cell by
 * x-eighth (fluid) with the wall particles last, and skip the eighths of the
 * stencil's outer cells that lie beyond 2h.  Every array a caller reads
 * (hash, index, cell ranges, sorted rows, measures in reference order, forces)
 * keeps its meaning; density and force sums run in another order, so they
 * agree with the plain walk to rounding, not bit for bit.  On by default for
 * the systems that qualify (DamBreak, DIV_CELL cells, no periodic images,
 * more than 2*cellcount+1 cells along x, cellcount * cell_size.x >= 2h);
 * CMAG_E_INVALID for others. */
int cmag_system_set_subcell_walk(cmag_system* s, int on);
int cmag_system_subcell_walk(const cmag_system* s);
/* In-range masks of the sub-cell walk (no counterpart in the reference): with
 * on != 0 (the default) the density pass leaves, for the blocks whose cell
 * lookups it stages, a 64-bit in-range mask per target and stencil row, and
 * the force pass takes the in-range neighbours of wall-free rows from it
 * instead of testing the candidates again.  Forces are the same bit for bit
 * either way; 72 bytes of device memory per particle. */
int cmag_system_set_inrange_masks(cmag_system* s, int on);
int cmag_system_inrange_masks(const cmag_system* s);
int cmag_system_synchronize(cmag_system* s);

/* stage timing (no reference counterpart): with timing on, update() records
 * HIP events between its stages on the object's stream; stage_times returns
 * the accumulated milliseconds and launch counts of each stage, in the order
 * hash, sort, reorder, density, force, integrate, and clears them. */
enum cmag_stage {
    CMAG_STAGE_HASH = 0, CMAG_STAGE_SORT = 1, CMAG_STAGE_REORDER = 2,
    CMAG_STAGE_DENSITY = 3, CMAG_STAGE_FORCE = 4, CMAG_STAGE_INTEGRATE = 5,
    CMAG_NUM_STAGES = 6
};
int cmag_system_enable_timing(cmag_system* s, int on);
int cmag_system_stage_times(cmag_system* s, double* ms, uint32_t* counts, int max_stages);

/* scenario controls: fluidSystem.cpp:222-255, poiseuilleFlowSystem.cpp:207-210 */
int cmag_system_change_right_boundary(cmag_system* s);
int cmag_system_remove_right_boundary(cmag_system* s);
int cmag_system_change_gravity(cmag_system* s);

/* ---- visualisation without OpenGL ------------------------------------------
 * The colour buffer the reference fills for its GL colour VBO (float4 RGBA per
 * particle, unsorted order), owned by the System, made on first call and
 * refilled by reset(): DamBreakSystem::_initialize fluidSystem.cpp:160-181
 * (colorRamp of 0.5 for the fluid, 0.2 for the first-type walls, 0.7 for the
 * rest; colorRamp :97-114), PoiseuilleFlowSystem::_initialize
 * poiseuilleFlowSystem.cpp:148-169, and for PeristalsisSystem its Coloring().
 * The device pointer is borrowed (getCudaColorVBO, fluidSystem.h:42). */
int cmag_system_color_buffer(cmag_system* s, void** dev_color4);
/* PeristalsisSystem::Coloring, PeristalsisSystem.cpp:409-442: walls black,
 * fluid left of worldOrigin.x + 4 radius red (255, 0, 0), the rest white
 * (255, 255, 255), from the current positions; CMAG_E_INVALID for the other
 * systems, which have no Coloring(). */
int cmag_system_coloring(cmag_system* s);
/* the same rules as a stage launcher, by the parameter block's system kind */
int cmag_fill_colors(const cmag_params* p, const float* pos4, float* color4, uint32_t n, cmag_stream_t stream);

/* The demos' particle renderer (DamBreak.Demo/render_particles.cpp:84-96 and
 * the sphere pixel shader of shaders.cpp, with the camera of main.cpp:19,
 * 109-114, 137), drawn into an RGBA8 image in device memory instead of a GL
 * window: every particle is a sprite of point_radius * (h / tan(fov / 2)) /
 * (eye distance) pixels (at most max_point_size), each pixel inside its disc
 * shaded colour * max(0, L . N), L = (0.577, 0.577, 0.577); the nearest
 * sprite wins a pixel (equal depths: the lower particle index, as GL_LESS in
 * draw order); a particle whose centre lies outside the view volume is not
 * drawn.  The image is stored top row first. */
typedef struct cmag_render_view {
    float translate[3];   /* glTranslatef (the demo's camera_trans) */
    float rotate[2];      /* degrees about x, then about y (camera_rot) */
    float fov_deg;        /* gluPerspective fovy (and setFOV) */
    float znear, zfar;
    float point_radius;   /* setradius: world-space sprite radius */
    float max_point_size; /* largest sprite side in pixels, 1..256 (0: 64) */
    float clear[4];       /* glClearColor */
} cmag_render_view;
typedef struct cmag_renderer cmag_renderer;
/* the demo's defaults: camera (0.5, 0, -2.5), no rotation, 60 degrees,
 * near 0.001, far 100, radius 0.0625, white background */
void cmag_render_view_default(cmag_render_view* v);
int cmag_renderer_create(uint32_t width, uint32_t height, cmag_renderer** out);
int cmag_renderer_destroy(cmag_renderer* r);
/* one frame of n particles (pos4: device float4 rows, color4: device float4
 * RGBA or NULL for white), enqueued on stream */
int cmag_renderer_display(cmag_renderer* r, const cmag_render_view* view, const float* pos4, const float* color4,
                          uint32_t n, cmag_stream_t stream);
/* the device image (width * height RGBA8, borrowed) and a synchronous copy of it to host memory */
void* cmag_renderer_image(cmag_renderer* r);
int cmag_renderer_read(cmag_renderer* r, uint8_t* rgba, cmag_stream_t stream);

/* ---- adjoint of the Poiseuille step (no reference counterpart) -------------
 * Vector-Jacobian products of this library's forward passes with the step's
 * neighbour structure held fixed (it is piecewise constant); walls (w == 1)
 * are not state.  They differentiate
 *   density + EOS     calculatePoiseuilleDensity, poiseuilleFlowSystem.cu:167-204
 *   force             calculatePoiseuilleAcceleration, poiseuilleFlowSystem.cu:206-247
 *   integrate         integratePoiseuilleSystem, poiseuilleFlowSystem.cu:84-101
 * Stage launchers (sorted order unless noted, float4 rows):
 *   cmag_integrate_vjp  unsorted; gpos/gvel/gvelLF: adjoints of the step's
 *                       outputs in, of its inputs out; gacc out; ggrav (FP64,
 *                       4 per particle, optional) += body-force adjoint shares
 *   cmag_force_vjp      from gaccS: gposS, gvelS (leap-frog velocity), grho
 *                       (one float per particle), gmu (optional, one float per
 *                       particle: the target's share of d/d mu), all written
 *   cmag_density_vjp    gposS += the density pass's share for grho */
int cmag_integrate_vjp(const cmag_params* p, const float* pos, float* gpos, float* gvel, float* gvelLF,
                       float* gacc, double* ggrav, uint32_t n, cmag_stream_t stream);
int cmag_force_vjp(const cmag_params* p, float* gposS, float* gvelS, float* grho, float* gmu, const float* gaccS,
                   const float* measures, const float* sortedPos, const float* sortedVel,
                   const uint32_t* cellStart, const uint32_t* cellEnd, uint32_t n, cmag_stream_t stream);
int cmag_density_vjp(const cmag_params* p, float* gposS, const float* grho, const float* sortedPos,
                     const uint32_t* cellStart, const uint32_t* cellEnd, uint32_t n, cmag_stream_t stream);
/* adjoint of one update() at the system's current state, which it leaves
 * unchanged (the neighbour structure and densities are recomputed):
 * gpos/gvel/gvelLF (device, unsorted, n float4 each) hold the adjoints of the
 * step's outputs on entry and of its inputs on return; gmu_acc (n doubles)
 * and ggrav_acc (4n doubles), both optional, accumulate the adjoints of mu
 * and of the body force per particle (sum them for the totals) */
int cmag_system_vjp_step(cmag_system* s, float* gpos, float* gvel, float* gvelLF, double* gmu_acc,
                         double* ggrav_acc);
/* device-to-device copy of rows [start, start+count) between a System array
 * and device memory (to_system != 0: into the System), ordered on the
 * System's stream */
int cmag_system_copy_array_device(cmag_system* s, int array, float* device_data, int start, int count,
                                  int to_system);

/* ---- batches: many independent systems in one step ------------------------
 * No reference counterpart: the reference's process-global __constant__
 * parameters (fluid_kernel.cu:17, fluidSystem.cu:17-19) allow one system per
 * process.  A batch holds m instances (any system kind; one dimension and hash
 * order per batch) with their own parameter blocks and particle counts in one
 * arena; cmag_batch_update steps every instance once, with the same per-particle
 * arithmetic as cmag_system_update on each instance alone (bit for bit).
 * Particles of instance k are numbered [first, first + count) of the arena
 * (cmag_batch_instance_range); set/get_array move one instance's float4 rows. */
typedef struct cmag_batch cmag_batch;

/* launch shape picked by cmag_batch_autotune (or set by hand):
 *   block            threads per workgroup and particles per tile: 64, 128, 256
 *   nb_cap           per-lane neighbour-list depth of the force pass (LDS): 8, 16, 32
 *   blocks_per_cu    persistent workgroups per CU, each running a contiguous
 *                    chunk of tiles; 0 = one workgroup per tile
 *   steps_per_graph  steps captured per HIP graph by cmag_batch_run */
typedef struct cmag_batch_tuning {
    uint32_t block;
    uint32_t nb_cap;
    uint32_t blocks_per_cu;
    uint32_t steps_per_graph;
} cmag_batch_tuning;

int cmag_batch_create(const cmag_params* params, const uint32_t* counts, uint32_t m, cmag_batch** out);
int cmag_batch_destroy(cmag_batch* b);
uint32_t cmag_batch_num_instances(const cmag_batch* b);
uint32_t cmag_batch_num_particles(const cmag_batch* b);
uint32_t cmag_batch_num_cells(const cmag_batch* b);
int cmag_batch_instance_range(const cmag_batch* b, uint32_t inst, uint32_t* first, uint32_t* count);
int cmag_batch_get_params(const cmag_batch* b, uint32_t inst, cmag_params* out);
/* count must equal the instance's particle count */
int cmag_batch_set_array(cmag_batch* b, uint32_t inst, int array, const float* data, uint32_t count);
int cmag_batch_get_array(cmag_batch* b, uint32_t inst, int array, float* data, uint32_t count);
void* cmag_batch_device_ptr(cmag_batch* b, int array);
/* elapsed time of one instance (kept on the device, read back) */
int cmag_batch_elapsed_time(cmag_batch* b, uint32_t inst, float* out);
/* elapsed time 0 and a fresh Peristalsis boundary-configuration state */
int cmag_batch_reset_clocks(cmag_batch* b);
/* one step of every instance: update() of each System, in one launch sequence */
int cmag_batch_update(cmag_batch* b);
/* steps steps; use_graph replays HIP graphs of steps_per_graph steps */
int cmag_batch_run(cmag_batch* b, uint32_t steps, int use_graph);
int cmag_batch_synchronize(cmag_batch* b);
int cmag_batch_get_tuning(const cmag_batch* b, cmag_batch_tuning* out);
int cmag_batch_set_tuning(cmag_batch* b, const cmag_batch_tuning* t);
/* times steps_per_trial steps of every launch shape on the batch's own state,
 * puts the state back after each trial, and keeps the fastest shape */
int cmag_batch_autotune(cmag_batch* b, uint32_t steps_per_trial, cmag_batch_tuning* best,
                        float* best_ms_per_step);
int cmag_batch_enable_timing(cmag_batch* b, int on);
int cmag_batch_stage_times(cmag_batch* b, double* ms, uint32_t* counts, int max_stages);

#ifdef __cplusplus
}
#endif
#endif /* CMAG_HIP_H */
