/*
 * cmag_params.h - plain-C parameter block for the WCSPH timestep.
 *
 * One POD struct replaces the three per-system parameter structs of the
 * reference, which were uploaded into a process-global __constant__ symbol:
 *   SimParams        DamBreak.Core/fluid_kernel.cuh:25-51
 *   PoiseuilleParams Poiseuille.Core/poiseuilleFlowKernel.cuh:16-35
 *   Peristalsiscfg   Peristalsis.Core/PeristalsisKernel.cuh:22-63
 * Here the block is passed to every kernel by value (kernel argument), so
 * several instances with different parameters can live in one process and in
 * one batched launch.  Every field is 4 bytes wide (no bool, no padding), so
 * the layout is identical for gcc, hipcc and Python ctypes.
 *
 * Derived fields (kernel coefficients, wall positions, clamp borders) are
 * precomputed on the host with the same float operation order the reference
 * used inside its kernels, so device and host see bit-identical constants.
 */
#ifndef CMAG_PARAMS_H
#define CMAG_PARAMS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CMAG_PARAMS_VERSION 1u

/* which reference system's loop bodies a step follows */
enum cmag_system_kind {
    CMAG_SYS_DAMBREAK    = 0, /* DamBreak.Core    (the north-star "FluidSystem") */
    CMAG_SYS_POISEUILLE  = 1, /* Poiseuille.Core */
    CMAG_SYS_PERISTALSIS = 2  /* Peristalsis.Core */
};

/* equation of state p(rho) */
enum cmag_eos {
    CMAG_EOS_TAIT      = 0, /* B((rho/rho0)^gamma - 1), fluid_kernel.cu:220 */
    CMAG_EOS_LINEAR_C2 = 1, /* c^2 rho,                 poiseuilleFlowKernel.cu:181 */
    CMAG_EOS_LINEAR_50C = 2 /* 50 c rho,                PeristalsisDensity.cu:75 */
};

/* cell index -> hash key */
enum cmag_hash_order {
    CMAG_HASH_LINEAR = 0, /* (z*gy)*gx + y*gx + x, fluid_kernel.cu:27-32 */
    CMAG_HASH_MORTON = 1  /* Z order: low morton_xbits bits of x, then x,y,z interleaved */
};

/* position -> cell coordinate rounding rule (bit-exactness depends on it) */
enum cmag_gridpos_rule {
    CMAG_GRIDPOS_DIV_CELL        = 0, /* floor((p-o)/cellSize),       fluid_kernel.cu:19-25 */
    CMAG_GRIDPOS_HALF_DIV_RADIUS = 1  /* floor(((p-o)*0.5f)/radius),  poiseuilleFlowKernel.cu:18-24 */
};

/* particle type codes kept in pos.w, as in the reference */
#define CMAG_DB_RIGHT_FIRST   0.0f /* fluid_kernel.cuh:18 */
#define CMAG_DB_FIRST         1.0f
#define CMAG_DB_RIGHT_SECOND  2.0f
#define CMAG_DB_SECOND        3.0f
#define CMAG_DB_FLUID         4.0f
#define CMAG_PO_FLUID         0.0f /* poiseuilleFlowSystem.cpp:340 */
#define CMAG_PO_WALL          1.0f
#define CMAG_PE_FLUID         0.0f /* PeristalsisSystem.cpp:298 walls: +-2,+-4,+-6 */

#define CMAG_CELL_EMPTY 0xffffffffu

typedef struct cmag_params {
    uint32_t version;          /* CMAG_PARAMS_VERSION */
    int32_t  system;           /* enum cmag_system_kind */
    int32_t  eos;              /* enum cmag_eos */
    int32_t  hash_order;       /* enum cmag_hash_order */
    int32_t  gridpos_rule;     /* enum cmag_gridpos_rule */
    int32_t  dim;              /* 2: every particle in one z cell, z stencil degenerate; 3: full */
    int32_t  periodic_mask;    /* bit a set: image shift on axis a (x=1,y=2,z=4) */
    int32_t  cellcount;        /* stencil half-width in cells */

    uint32_t grid_size[3];     /* cells per axis, each a power of two */
    uint32_t grid_bits[3];     /* log2(grid_size) */
    uint32_t fluid_size[3];    /* fluid lattice extent */
    float    world_origin[3];
    float    cell_size[3];
    float    world_size[3];    /* periodic image length per axis */

    float    gravity[3];
    float    particle_radius;  /* r (lattice spacing is 2r) */
    float    smoothing_radius; /* h */
    float    particle_mass;
    float    rest_density;
    float    soundspeed;
    float    B;                /* Tait stiffness */
    float    gamma;            /* Tait exponent (integer valued, <= 15) */
    float    delta_time;
    float    boundary_damping;
    float    mu;               /* dynamic viscosity (Morris term) */
    float    alpha;            /* artificial viscosity coefficient (0.38 in the reference) */
    float    kernel_coeff;     /* W(q) normalisation, 7/(4 pi h^2) in 2-D */
    float    grad_coeff;       /* dW/dr normalisation, 7/(2 pi h^3) in 2-D */

    int32_t  boundary_offset;
    float    right_boundary;   /* DamBreak right clamp (moves with changeRightBoundary) */
    float    lj_D;             /* Lennard-Jones wall strength */
    float    lj_a;             /* Lennard-Jones length */
    float    clamp_lo[3];      /* DamBreak integrate clamps; -inf disables */
    float    clamp_hi[3];      /* x uses right_boundary; +inf disables */

    float    wall_bottom;      /* Poiseuille Morris extrapolation planes */
    float    wall_top;
    float    beta_max;         /* 1000 in the reference */

    float    amplitude;        /* Peristalsis wave amplitude */
    float    wave_speed;
    float    wave_k;           /* wavenumber: pi/(fluid_size.x*radius) in the reference */
    int32_t  is_boundary_configuration;
    int32_t  wall_velocity;    /* 1: moving walls carry their kinematic velocity */

    uint32_t morton_mask[3];   /* key bits owned by each axis (Morton order) */
    uint32_t morton_xbits;     /* Morton order over 2^xbits x 1 x 1 cell bricks (0 = plain) */
    int32_t  reserved[4];
} cmag_params;

#ifdef __cplusplus
}
#endif

#endif /* CMAG_PARAMS_H */
