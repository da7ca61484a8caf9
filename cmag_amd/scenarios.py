"""Parameter blocks and initial states.

Reference blocks come from the library's constructors (cmag_params_*), which
restate the reference constructors.  Configuration scenarios (the channel and
dam-break workloads this package is measured on) are built here:

  poiseuille_channel  2-D periodic channel, Tait EOS, 3 wall layers, seeded jitter
  dambreak3d          3-D water column in a tank with 3-layer walls, open top

Particle mass in configuration mode is rho0 over the lattice kernel sum (the
method of PeristalsisSystem::CalculateMass, PeristalsisSystem.cpp:238-253,
taken over an unbounded lattice), so an interior particle of the unperturbed
lattice has density rho0.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from ._lib import CmagParams, call

PI_F = np.float32(3.141592654)


# ---------------------------------------------------------------- reference
def params_dambreak(fluid, bo, grid, r) -> CmagParams:
    p = CmagParams()
    call("cmag_params_dambreak", C.byref(p), *map(int, fluid), int(bo), *map(int, grid), C.c_float(r))
    return p


def params_poiseuille(fluid, bo, grid, r) -> CmagParams:
    p = CmagParams()
    call("cmag_params_poiseuille", C.byref(p), *map(int, fluid), int(bo), *map(int, grid), C.c_float(r))
    return p


def params_peristalsis(dt, fluid, amplitude, wave_speed, soundspeed, gravity, bo, grid, r) -> CmagParams:
    p = CmagParams()
    call("cmag_params_peristalsis", C.byref(p), C.c_float(dt), *map(int, fluid), C.c_float(amplitude),
         C.c_float(wave_speed), C.c_float(soundspeed), *[C.c_float(float(g)) for g in gravity], int(bo),
         *map(int, grid), C.c_float(r))
    return p


def finalize(p: CmagParams) -> CmagParams:
    call("cmag_params_finalize", C.byref(p))
    return p


def num_cells(p: CmagParams) -> int:
    return int(p.grid_size[0]) * int(p.grid_size[1]) * int(p.grid_size[2])


def key_bits(p: CmagParams) -> int:
    return int(p.grid_bits[0] + p.grid_bits[1] + p.grid_bits[2])


# ---------------------------------------------------------------- utilities
class SplitMix64:
    """splitmix64 (Steele, Lea, Flood 2014); uniform floats from the top 24 bits."""

    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.state = seed & self.MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & self.MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self, n: int) -> np.ndarray:
        """n doubles in [0,1) (vectorised splitmix64 over consecutive states)."""
        k = np.arange(1, n + 1, dtype=np.uint64)
        with np.errstate(over="ignore"):
            z = np.uint64(self.state) + k * np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        self.state = (self.state + n * 0x9E3779B97F4A7C15) & self.MASK
        return (z >> np.uint64(40)).astype(np.float64) / float(1 << 24)


def wendland_w(q: np.ndarray, h: float, dim: int) -> np.ndarray:
    """Wendland C2 kernel W(q) with the 2-D 7/(4 pi h^2) or 3-D 21/(16 pi h^3) factor."""
    a = 7.0 / (4.0 * math.pi * h * h) if dim == 2 else 21.0 / (16.0 * math.pi * h ** 3)
    q = np.asarray(q, dtype=np.float64)
    return np.where(q < 2.0, a * (1.0 - 0.5 * q) ** 4 * (2.0 * q + 1.0), 0.0)


def lattice_kernel_sum(h: float, dx: float, dim: int) -> float:
    """sum_j W(|x_j|) over an unbounded square/cubic lattice of spacing dx (FP64)."""
    k = int(math.ceil(2.0 * h / dx)) + 1
    r = np.arange(-k, k + 1, dtype=np.float64) * dx
    if dim == 2:
        X, Y = np.meshgrid(r, r, indexing="ij")
        d = np.sqrt(X * X + Y * Y)
    else:
        X, Y, Z = np.meshgrid(r, r, r, indexing="ij")
        d = np.sqrt(X * X + Y * Y + Z * Z)
    return float(wendland_w(d / h, h, dim).sum())


def lattice_laplacian_moment(h: float, dx: float, dim: int) -> float:
    """-(dx^dim / dim) sum_j |x_j| W'(|x_j|) over the lattice: the factor by
    which the Morris viscous sum reproduces nu * laplacian(u) on that lattice
    (1 in the continuum limit; 0.974 at h = 1.3 dx in 2-D)."""
    k = int(math.ceil(2.0 * h / dx)) + 1
    r = np.arange(-k, k + 1, dtype=np.float64) * dx
    grids = np.meshgrid(*([r] * dim), indexing="ij")
    d = np.sqrt(sum(g * g for g in grids))
    q = d / h
    a = 7.0 / (2.0 * math.pi * h ** 3) if dim == 2 else 21.0 / (8.0 * math.pi * h ** 4)
    t = np.clip(1.0 - 0.5 * q, 0.0, None)
    dw = np.where(q < 2.0, a * (t ** 4 - t ** 3 * (2.0 * q + 1.0)), 0.0)
    with np.errstate(invalid="ignore", divide="ignore"):
        term = np.where(d > 0, d * dw, 0.0)
    return float(-term.sum() * dx ** dim / dim)


def _pow2_at_least(v: float) -> int:
    n = 1
    while n < v:
        n *= 2
    return n


def _blank(system: int, dim: int) -> CmagParams:
    p = CmagParams()
    p.version = 1
    p.system = system
    p.dim = dim
    p.hash_order = L.HASH_LINEAR
    p.gridpos_rule = L.GRIDPOS_DIV_CELL
    p.boundary_damping = -1.0
    p.alpha = 0.38
    p.beta_max = 1000.0
    p.gamma = 7.0
    p.rest_density = 1000.0
    inf = float("inf")
    for a in range(3):
        p.clamp_lo[a] = -inf
        p.clamp_hi[a] = inf
    p.right_boundary = inf
    return p


# ---------------------------------------------------------------- channel
@dataclass
class Scenario:
    params: CmagParams
    pos: np.ndarray        # (N, 4) float32, w = particle type
    vel: np.ndarray        # (N, 4) float32
    num_fluid: int
    meta: dict


def poiseuille_channel(nx: int = 32, ny: int = 64, width: float = 1.0, bo: int = 3,
                       nu: float = 1e-2, vmax: float = 1e-2, h_factor: float = 1.3,
                       jitter: float = 1e-3, seed: int = 1973, eos: int = L.EOS_TAIT,
                       cellcount: int | None = None, cell_factor: float = 1.0,
                       gz: int = 4, gravity: float | None = None) -> Scenario:
    """2-D periodic Poiseuille channel in the Poiseuille system's loop bodies.

    Fluid nx x ny at spacing dx = width/ny, periodic in x with length nx*dx;
    bo wall layers (w = 1) above and below.  Body force g = 8 nu vmax / width^2,
    sound speed c0 = 10 vmax, Tait gamma 7, dt = min(0.25 h/c0, 0.125 h^2/nu).
    Cells are cell_factor*dx wide along x (nx/cell_factor must be a power of
    two so the grid spans the period exactly)."""
    dx = width / ny
    r = dx / 2.0
    h = h_factor * dx
    cs = cell_factor * dx
    gx = int(round(nx / cell_factor))
    if gx * cell_factor != nx or gx & (gx - 1):
        raise ValueError("nx / cell_factor must be a power of two")
    if cellcount is None:
        cellcount = int(math.ceil(2.0 * h / cs - 1e-9))
    height = (ny + 2 * bo) * dx
    gy = _pow2_at_least(height / cs + 2)
    p = _blank(L.SYS_POISEUILLE, 2)
    p.eos = eos
    p.periodic_mask = 1
    p.cellcount = cellcount
    for a, g in enumerate((gx, gy, gz)):
        p.grid_size[a] = g
    p.fluid_size[0], p.fluid_size[1], p.fluid_size[2] = nx, ny, 1
    p.particle_radius = r
    p.smoothing_radius = h
    p.cell_size[0] = p.cell_size[1] = p.cell_size[2] = cs
    # world: x periodic over [-nx dx/2, nx dx/2); y from the bottom of the walls
    p.world_origin[0] = np.float32(-0.5 * nx * dx)
    p.world_origin[1] = np.float32(-0.5 * gy * cs)
    p.world_origin[2] = np.float32(-0.5 * gz * cs)
    p.world_size[0] = np.float32(gx) * np.float32(cs)
    p.world_size[1] = np.float32(gy) * np.float32(cs)
    p.world_size[2] = np.float32(gz) * np.float32(cs)
    c0 = 10.0 * vmax
    p.soundspeed = c0
    p.B = 1000.0 * c0 * c0 / 7.0
    p.mu = 1000.0 * nu
    g = 8.0 * nu * vmax / (width * width) if gravity is None else gravity
    p.gravity[0], p.gravity[1], p.gravity[2] = g, 0.0, 0.0
    p.delta_time = min(0.25 * h / c0, 0.125 * h * h / nu)
    p.particle_mass = 1000.0 / lattice_kernel_sum(h, dx, 2)
    # channel: fluid rows y0 + (j+1/2) dx for j in [0, ny); walls below/above
    y0 = np.float32(-0.5 * height + bo * dx)
    p.wall_bottom = y0
    p.wall_top = np.float32(y0 + ny * dx)
    finalize(p)

    rng = SplitMix64(seed)
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    fx = (np.float64(p.world_origin[0]) + (ii.ravel() + 0.5) * dx)
    fy = (np.float64(y0) + (jj.ravel() + 0.5) * dx)
    nf = nx * ny
    jit = (rng.uniform(2 * nf).reshape(2, nf) * 2.0 - 1.0) * jitter * dx
    fluid = np.zeros((nf, 4), np.float64)
    fluid[:, 0] = fx + jit[0]
    fluid[:, 1] = fy + jit[1]
    fluid[:, 2] = float(p.world_origin[2]) + 0.5 * cs
    wi, wj = np.meshgrid(np.arange(nx), np.arange(bo), indexing="xy")
    wx = np.float64(p.world_origin[0]) + (wi.ravel() + 0.5) * dx
    bottom = np.zeros((nx * bo, 4)); top = np.zeros((nx * bo, 4))
    bottom[:, 0] = wx
    bottom[:, 1] = np.float64(y0) - (wj.ravel() + 0.5) * dx
    top[:, 0] = wx
    top[:, 1] = np.float64(y0) + ny * dx + (wj.ravel() + 0.5) * dx
    bottom[:, 2] = top[:, 2] = fluid[0, 2]
    bottom[:, 3] = top[:, 3] = 1.0
    pos = np.concatenate([fluid, bottom, top]).astype(np.float32)
    vel = np.zeros_like(pos)
    meta = dict(dx=dx, h=h, nu=nu, vmax=vmax, width=width, g=g, y0=float(y0), nx=nx, ny=ny)
    return Scenario(p, pos, vel, nf, meta)


def poiseuille_series(y: np.ndarray, t: float, F: float, l: float, nu: float,
                      nterms: int = 101) -> np.ndarray:
    """Transient start-up Poiseuille profile, the series of
    Poiseuille.Report/XVelocityYPosition.m:12 (the script plots -Vel; this is the
    positive velocity):  u = F y (l-y)/(2 nu) - sum_n 4 F l^2 sin((2n+1) pi y/l)
    exp(-(2n+1)^2 pi^2 nu t / l^2) / (nu pi^3 (2n+1)^3)."""
    y = np.asarray(y, dtype=np.float64)
    u = F * y * (l - y) / (2.0 * nu)
    for n in range(nterms):
        k = 2 * n + 1
        u -= (4.0 * F * l * l * np.sin(k * math.pi * y / l) * math.exp(-k * k * math.pi ** 2 * nu * t / (l * l))
              / (nu * math.pi ** 3 * k ** 3))
    return u


# ---------------------------------------------------------------- dam break
def dambreak3d(dx: float = 5e-3, tank=(1.6, 1.0, 0.4), column=(0.4, 0.8, 0.4), layers: int = 3,
               h_factor: float = 1.3, alpha: float = 0.01, jitter: float = 0.01, seed: int = 1973,
               grid=None, hash_order: int = L.HASH_LINEAR, cell_factor: float = 1.0,
               morton_xbits: int = 0, gate: bool | None = None) -> Scenario:
    """3-D dam break in the DamBreak system's loop bodies (Tait EOS, Monaghan
    artificial viscosity, Lennard-Jones wall particles, clamp walls).

    Water column column[0] x column[1] x column[2] against the x=0 wall of a
    tank tank[0] x tank[1] x tank[2]; `layers` wall layers on the floor and the
    four sides (open top); lattice spacing dx with seeded jitter +-jitter*dx.
    Cells are cell_factor * 2h wide (cellcount = ceil(2h / cell)); with the
    Morton hash, morton_xbits > 0 orders bricks of 2^xbits cells along x.

    gate: the tank's right wall starts at the column's face as a gate of
    movable wall particles (type RightFirst, as DamBreak.Core/fluidSystem.cpp's
    right boundary) with the clamp right_boundary there too, so the column can
    relax before it is released; changeRightBoundary moves gate and clamp by
    the column width (fluid_size.x * 2r) per call, and
    meta["release_shifts"] calls put both at the tank's far wall (release).
    Without the gate the column starts released.  gate=None: gated where the
    lattice allows it (tank width a whole multiple of the column width in
    lattice steps), else released."""
    h = h_factor * dx
    cs = cell_factor * 2.0 * h
    cellcount = int(math.ceil(2.0 * h / cs - 1e-9))
    nfx, nfy, nfz = (int(round(c / dx)) for c in column)
    ext = [t + 2 * layers * dx for t in tank]
    if grid is None:
        grid = tuple(_pow2_at_least(e / cs + 1) for e in ext)
    p = _blank(L.SYS_DAMBREAK, 3)
    p.hash_order = hash_order
    p.morton_xbits = morton_xbits
    p.eos = L.EOS_TAIT
    p.cellcount = cellcount
    for a in range(3):
        p.grid_size[a] = grid[a]
        p.cell_size[a] = cs
        p.world_origin[a] = np.float32(-layers * dx - 0.5 * (grid[a] * cs - ext[a]))
        p.world_size[a] = np.float32(grid[a] * cs)
    p.fluid_size[0], p.fluid_size[1], p.fluid_size[2] = nfx, nfy, nfz
    p.particle_radius = dx / 2.0
    p.smoothing_radius = h
    g = 9.81
    H = column[1]
    c0 = 10.0 * math.sqrt(2.0 * g * H)
    p.soundspeed = c0
    p.gamma = 7.0
    p.B = 1000.0 * c0 * c0 / 7.0
    p.alpha = alpha
    p.gravity[0], p.gravity[1], p.gravity[2] = 0.0, -g, 0.0
    p.delta_time = 0.25 * h / c0
    p.particle_mass = 1000.0 / lattice_kernel_sum(h, dx, 3)
    # Lennard-Jones walls scaled as the reference scales them (fluidSystem.cpp:70-71)
    p.lj_D = 10.0 * (-g) * H
    p.lj_a = dx / 2.0
    # clamp box = tank interior
    p.clamp_lo[0], p.clamp_lo[1], p.clamp_lo[2] = 0.0, 0.0, 0.0
    p.right_boundary = tank[0]
    p.clamp_hi[1] = float("inf")
    p.clamp_hi[2] = tank[2]
    p.boundary_offset = layers
    finalize(p)

    rng = SplitMix64(seed)
    I, J, K = np.meshgrid(np.arange(nfx), np.arange(nfy), np.arange(nfz), indexing="ij")
    nf = I.size
    # lattice order x fastest, then y, then z (as initFluid)
    I = I.transpose(2, 1, 0).ravel(); J = J.transpose(2, 1, 0).ravel(); K = K.transpose(2, 1, 0).ravel()
    jit = (rng.uniform(3 * nf).reshape(3, nf) * 2.0 - 1.0) * jitter * dx
    fluid = np.empty((nf, 4), np.float64)
    fluid[:, 0] = (I + 0.5) * dx + jit[0]
    fluid[:, 1] = (J + 0.5) * dx + jit[1]
    fluid[:, 2] = (K + 0.5) * dx + jit[2]
    fluid[:, 3] = 4.0  # DamBreak fluid type
    # walls: floor and four sides, `layers` deep, spacing dx, outside the tank
    nx_t, ny_t, nz_t = (int(round(t / dx)) for t in tank)
    xs = (np.arange(-layers, nx_t + layers) + 0.5) * dx
    zs = (np.arange(-layers, nz_t + layers) + 0.5) * dx
    ys = (np.arange(0, ny_t) + 0.5) * dx
    walls = []
    for k in range(layers):
        y = -(k + 0.5) * dx
        X, Z = np.meshgrid(xs, zs, indexing="ij")
        walls.append(np.stack([X.ravel(), np.full(X.size, y), Z.ravel()], 1))
    left_x = (-np.arange(1, layers + 1) + 0.5) * dx
    right_x = (nx_t + np.arange(layers) + 0.5) * dx
    side_z = np.concatenate([(-np.arange(1, layers + 1) + 0.5) * dx, (nz_t + np.arange(layers) + 0.5) * dx])
    X, Y, Z = np.meshgrid(left_x, ys, zs, indexing="ij")
    walls.append(np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1))
    inner_z = (np.arange(0, nz_t) + 0.5) * dx
    fits = nfx > 0 and nx_t % nfx == 0 and nx_t > nfx
    if gate is None:
        gate = fits
    elif gate and not fits:
        raise ValueError("the gate reaches the tank's far wall only if the tank is a whole multiple of the column "
                         "(in lattice steps)")
    if gate:
        # the right wall's columns beside the tank's z walls stay put; its
        # inner part is the gate, at the column's face until released
        X, Y, Z = np.meshgrid(right_x, ys, side_z, indexing="ij")
        walls.append(np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1))
    else:
        X, Y, Z = np.meshgrid(right_x, ys, zs, indexing="ij")
        walls.append(np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1))
    inner_x = (np.arange(0, nx_t) + 0.5) * dx
    X, Y, Z = np.meshgrid(inner_x, ys, side_z, indexing="ij")
    walls.append(np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1))
    w = np.concatenate(walls)
    wall = np.empty((w.shape[0], 4), np.float64)
    wall[:, :3] = w
    wall[:, 3] = 1.0  # DamBreak FirstType wall
    parts = [fluid, wall]
    shifts = 0
    if gate:
        shifts = nx_t // nfx - 1
        gx = right_x - shifts * nfx * dx
        X, Y, Z = np.meshgrid(gx, ys, inner_z, indexing="ij")
        g4 = np.empty((X.size, 4), np.float64)
        g4[:, 0], g4[:, 1], g4[:, 2] = X.ravel(), Y.ravel(), Z.ravel()
        g4[:, 3] = 0.0  # DamBreak RightFirstType: moved by changeRightBoundary
        parts.append(g4)
        p.right_boundary = np.float32(nfx * dx)
    pos = np.concatenate(parts).astype(np.float32)
    vel = np.zeros_like(pos)
    meta = dict(dx=dx, h=h, c0=c0, tank=tank, column=column, num_wall=int(pos.shape[0] - nf),
                release_shifts=shifts)
    return Scenario(p, pos, vel, nf, meta)


# Martin & Moyce (1952), collapse of a liquid column with n^2 = 2 (height
# twice the width a): surge front Z = x_front / a against T = t sqrt(2g/a)
# (their Table 2, as the reference's DamBreak.Report/XFrontTest.h uses it)
MARTIN_MOYCE_SURGE = ((0.00, 1.00), (0.41, 1.11), (0.84, 1.22), (1.19, 1.44), (1.43, 1.67), (1.63, 1.89),
                      (1.83, 2.11), (1.98, 2.33), (2.20, 2.56), (2.32, 2.78), (2.51, 3.00), (2.65, 3.22),
                      (2.83, 3.44), (2.97, 3.67), (3.11, 3.89))


def relax_gated(system, sc: Scenario, seconds: float = 0.3, alpha: float = 2.0, zero_every: float = 0.02,
                free_last: float = 0.02, report=None, use_graph: bool = True) -> None:
    """Let a gated dam break settle behind its gate: stepping with a strong
    artificial viscosity and the velocities set to zero every zero_every
    seconds (kinetic damping), the last free_last seconds undamped, so that
    what is left is the column's own residual motion.  The parameter block's
    alpha is restored afterwards; the clock keeps running.  report(done,
    total), if given, is called after every batch of steps; use_graph=False
    steps eagerly instead of replaying a captured step."""
    P0 = system.params
    P = P0.copy()
    P.alpha = alpha
    system.set_params(P)
    dt = float(P.delta_time)
    total = int(round(seconds / dt))
    every = max(1, int(round(zero_every / dt)))
    free = int(round(free_last / dt))
    zeros = np.zeros((system.getNumParticles(), 4), np.float32)
    k = 0
    while k < total:
        m = min(every, total - k)
        system.run(m, use_graph=use_graph)
        k += m
        if report is not None:
            report(k, total)
        if k < total - free:
            system.setArray(L.ARRAY_VELOCITY, zeros)
            system.setArray(L.ARRAY_VELOCITYLEAPFROG, zeros)
    system.set_params(P0)


def release_gate(system, sc: Scenario) -> None:
    """open the gate: move gate particles and clamp to the tank's far wall"""
    for _ in range(sc.meta["release_shifts"]):
        system.changeRightBoundary()


def dambreak3d_particles(target_fluid: int, **kw) -> Scenario:
    """Dam break with the lattice spacing chosen for about target_fluid fluid particles
    (0.4 x 0.8 x 0.4 column, dx = 0.4 / round((target/2)^(1/3)))."""
    n = int(round((target_fluid / 2.0) ** (1.0 / 3.0)))
    return dambreak3d(dx=0.4 / n, **kw)


# ---------------------------------------------------------------- channel batches
def _channel_grid(p: CmagParams, nx: int, dx: float, height: float, h: float, cell_factor: float) -> None:
    """x-periodic 2-D grid over a channel of nx lattice columns: gx cells span the
    period nx*dx exactly (gx the largest power of two with cell >= cell_factor*dx),
    square cells, gy covering `height` plus a margin, one cell along z."""
    period = nx * dx
    gx = 1
    while 2 * gx * cell_factor * dx <= period + 1e-12:
        gx *= 2
    cs = period / gx
    gy = _pow2_at_least(height / cs + 4)
    p.cellcount = int(math.ceil(2.0 * h / cs - 1e-9))
    p.grid_size[0], p.grid_size[1], p.grid_size[2] = gx, gy, 1
    p.cell_size[0] = p.cell_size[1] = p.cell_size[2] = cs
    p.world_origin[0] = np.float32(-0.5 * period)
    p.world_origin[1] = np.float32(-0.5 * gy * cs)
    p.world_origin[2] = np.float32(-0.5 * cs)
    p.world_size[0] = np.float32(gx) * np.float32(cs)
    p.world_size[1] = np.float32(gy) * np.float32(cs)
    p.world_size[2] = np.float32(cs)
    p.periodic_mask = 1


def channel_poiseuille(n_target: int, nu: float, seed: int, width: float = 1.0, bo: int = 3,
                       vmax: float = 1e-2, h_factor: float = 1.3, aspect: float = 0.5,
                       jitter: float = 1e-3, cell_factor: float = 1.0) -> Scenario:
    """One Poiseuille channel of about n_target particles (fluid + walls) in the
    Poiseuille system's loop bodies: fluid nx x ny with nx ~ aspect*ny, dx =
    width/ny, bo wall layers (w = 1) above and below, body force 8 nu vmax /
    width^2, Tait EOS with c0 = 10 vmax, dt = min(0.25 h/c0, 0.125 h^2/nu)."""
    ny = int(round((-2 * bo + math.sqrt(4 * bo * bo + 4 * n_target / aspect)) / 2))
    ny = max(8, ny)
    nx = max(4, int(round(aspect * ny)))
    dx = width / ny
    h = h_factor * dx
    p = _blank(L.SYS_POISEUILLE, 2)
    p.eos = L.EOS_TAIT
    _channel_grid(p, nx, dx, (ny + 2 * bo) * dx, h, cell_factor)
    p.fluid_size[0], p.fluid_size[1], p.fluid_size[2] = nx, ny, 1
    p.particle_radius = dx / 2.0
    p.smoothing_radius = h
    c0 = 10.0 * vmax
    p.soundspeed = c0
    p.B = 1000.0 * c0 * c0 / 7.0
    p.mu = 1000.0 * nu
    g = 8.0 * nu * vmax / (width * width)
    p.gravity[0] = g
    p.delta_time = min(0.25 * h / c0, 0.125 * h * h / nu)
    p.particle_mass = 1000.0 / lattice_kernel_sum(h, dx, 2)
    y0 = np.float32(-0.5 * ny * dx)
    p.wall_bottom = y0
    p.wall_top = np.float32(y0 + ny * dx)
    finalize(p)
    rng = SplitMix64(seed)
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    nf = nx * ny
    jit = (rng.uniform(2 * nf).reshape(2, nf) * 2.0 - 1.0) * jitter * dx
    z = float(p.world_origin[2]) + 0.5 * float(p.cell_size[2])
    fluid = np.zeros((nf, 4))
    fluid[:, 0] = float(p.world_origin[0]) + (ii.ravel() + 0.5) * dx + jit[0]
    fluid[:, 1] = float(y0) + (jj.ravel() + 0.5) * dx + jit[1]
    fluid[:, 2] = z
    wi, wj = np.meshgrid(np.arange(nx), np.arange(bo), indexing="xy")
    wx = float(p.world_origin[0]) + (wi.ravel() + 0.5) * dx
    walls = np.zeros((2 * nx * bo, 4))
    walls[:nx * bo, 0] = wx
    walls[:nx * bo, 1] = float(y0) - (wj.ravel() + 0.5) * dx
    walls[nx * bo:, 0] = wx
    walls[nx * bo:, 1] = float(y0) + ny * dx + (wj.ravel() + 0.5) * dx
    walls[:, 2] = z
    walls[:, 3] = 1.0
    pos = np.concatenate([fluid, walls]).astype(np.float32)
    meta = dict(kind="poiseuille", dx=dx, h=h, nu=nu, vmax=vmax, g=g, nx=nx, ny=ny, width=width)
    return Scenario(p, pos, np.zeros_like(pos), nf, meta)


def channel_peristalsis(n_target: int, amp_ratio: float, wl_ratio: float, seed: int, width: float = 1.0,
                        bo: int = 3, wave_speed: float = 1e-2, nu: float = 1e-2, h_factor: float = 1.5,
                        jitter: float = 1e-3, cell_factor: float = 1.0) -> Scenario:
    """One peristaltic channel of about n_target particles in the Peristalsis
    system's loop bodies (PeristalsisDensity.cu, PeristalsisPressureForce.cu,
    PeristalsisViscousForce.cu, PeristalsisIntegrate.cu): half-width H =
    width/2 = ny*r, walls w = +-2, +-4, +-6 on the sinusoid
    y = -+(H - A sin(k(x - x0) - k c t)) -+ layer offset with A = amp_ratio*H,
    wavelength = domain length = wl_ratio*width (k = pi/(nx r) as in GetWave,
    PeristalsisKernel.cuh:59-62), wave speed c.  The walls start on the
    sinusoid (no configuration ramp) and carry the wall velocity; the fluid is
    the lattice inside the channel at t = 0.  EOS p = 50 c_s rho with 50 c_s =
    (10 max(c, A k c))^2, mu = 1000 nu, no body force."""
    ny = int(round((-2 * bo + math.sqrt(4 * bo * bo + 4 * n_target / wl_ratio)) / 2))
    ny = max(8, ny)
    nx = max(4, int(round(wl_ratio * ny)))
    dx = width / ny
    r = dx / 2.0
    h = h_factor * dx
    H = ny * r
    A = amp_ratio * H
    p = _blank(L.SYS_PERISTALSIS, 2)
    p.eos = L.EOS_LINEAR_50C
    _channel_grid(p, nx, dx, 2 * (H + A) + 2 * bo * dx, h, cell_factor)
    p.fluid_size[0], p.fluid_size[1], p.fluid_size[2] = nx, ny, 1
    p.boundary_offset = bo
    p.particle_radius = r
    p.smoothing_radius = h
    p.amplitude = A
    p.wave_speed = wave_speed
    p.wave_k = np.float32(PI_F / np.float32(nx * np.float32(r)))
    k = 2.0 * math.pi / (nx * dx)
    c_eff = 10.0 * max(wave_speed, A * k * wave_speed)
    p.soundspeed = c_eff * c_eff / 50.0
    p.mu = 1000.0 * nu
    p.delta_time = min(0.25 * h / c_eff, 0.125 * h * h / nu)
    p.particle_mass = 1000.0 / lattice_kernel_sum(h, dx, 2)
    p.is_boundary_configuration = 0
    p.wall_velocity = 1
    finalize(p)
    ox = np.float32(p.world_origin[0])
    wk = np.float32(p.wave_k)
    rng = SplitMix64(seed)
    # fluid lattice over |y| < H + A, kept inside the channel at t = 0
    rows = int(math.ceil(2 * (H + A) / dx))
    ii, jj = np.meshgrid(np.arange(nx), np.arange(rows), indexing="xy")
    x = float(ox) + (ii.ravel() + 0.5) * dx
    y = -(H + A) + (jj.ravel() + 0.5) * dx
    s = np.sin(k * (x - float(ox)))
    keep = (y >= -H + A * s + 0.5 * dx) & (y <= H - A * s - 0.5 * dx)
    x, y = x[keep], y[keep]
    nf = x.size
    jit = (rng.uniform(2 * nf).reshape(2, nf) * 2.0 - 1.0) * jitter * dx
    z = float(p.world_origin[2]) + 0.5 * float(p.cell_size[2])
    fluid = np.zeros((nf, 4))
    fluid[:, 0] = x + jit[0]
    fluid[:, 1] = y + jit[1]
    fluid[:, 2] = z
    # walls in float32 with the kernels' formula (PeristalsisKernel.cu:144-171)
    wx = (ox + np.float32(r) * (2 * np.arange(nx, dtype=np.float32) + 1)).astype(np.float32)
    wave = np.sin((wk * (wx - ox)).astype(np.float32)).astype(np.float32)
    nyr = np.float32(ny) * np.float32(r)
    walls = []
    for layer in range(bo):
        wcode = np.float32(2.0 * (1 + layer))
        yb = -nyr + np.float32(A) * wave - np.float32(r) * (wcode - 1)
        yt = nyr - np.float32(A) * wave + np.float32(r) * (wcode - 1)
        for yy, code in ((yb, wcode), (yt, -wcode)):
            wl = np.zeros((nx, 4), np.float32)
            wl[:, 0] = wx
            wl[:, 1] = yy
            wl[:, 2] = z
            wl[:, 3] = code
            walls.append(wl)
    pos = np.concatenate([fluid.astype(np.float32)] + walls).astype(np.float32)
    meta = dict(kind="peristalsis", dx=dx, h=h, nu=nu, amplitude=A, wavelength=nx * dx, nx=nx, ny=ny,
                amp_ratio=amp_ratio, wl_ratio=wl_ratio, width=width)
    return Scenario(p, pos, np.zeros_like(pos), nf, meta)


def channel_batch(m: int = 256, seed: int = 20240, n_lo: int = 1024, n_hi: int = 65536,
                  nu_lo: float = 1e-3, nu_hi: float = 1e-1, cell_factor: float = 1.0) -> list[Scenario]:
    """The batched workload: m independent 2-D channels with particle counts
    log-uniform in [n_lo, n_hi]; the first half Poiseuille with nu log-uniform in
    [nu_lo, nu_hi], the second half peristaltic with amplitude ratio U(0.1, 0.5)
    and wavelength U(0.5, 2) x width.  Seeded (splitmix64)."""
    rng = SplitMix64(seed)
    out = []
    ln = math.log
    for k in range(m):
        u = rng.uniform(4)
        n = int(round(math.exp(ln(n_lo) + float(u[0]) * (ln(n_hi) - ln(n_lo)))))
        if k < m // 2:
            nu = math.exp(ln(nu_lo) + float(u[1]) * (ln(nu_hi) - ln(nu_lo)))
            out.append(channel_poiseuille(n, nu, seed=seed + k, cell_factor=cell_factor))
        else:
            amp = 0.1 + 0.4 * float(u[2])
            wl = 0.5 + 1.5 * float(u[3])
            out.append(channel_peristalsis(n, amp, wl, seed=seed + k, cell_factor=cell_factor))
    return out


# ---------------------------------------------------------------- adjoint workload
def adjoint_channel(seed: int = 1973, nu_star: float = 2e-2, nu0: float = 1e-2, g0_factor: float = 1.3,
                    vmax: float = 1e-2, nx: int = 128, ny: int = 128, steps: int = 200,
                    checkpoint_every: int = 20, bins: int = 64) -> dict:
    """The differentiable workload: a 2-D Poiseuille channel of nx x ny fluid
    particles (dx = 1/ny, x-period nx*dx), 3 wall layers, cfg-1 physics (Tait,
    c0 = 10 vmax, h = 1.3 dx).  The target profile comes from (nu*, g*) with
    g* = 8 nu* vmax (0.5 + U) (seeded); the gradient is taken at (nu0, g0 =
    g0_factor g*).  The time step is fixed for both, from the larger
    viscosity.  Returns the scenario, the two parameter blocks, and the run
    lengths."""
    rng = SplitMix64(seed)
    g_star = 8.0 * nu_star * vmax * (0.5 + float(rng.uniform(1)[0]))
    sc = poiseuille_channel(nx=nx, ny=ny, nu=max(nu_star, nu0), vmax=vmax, gravity=g_star, seed=seed)
    p_star = sc.params.copy()
    p_star.mu = 1000.0 * nu_star
    p0 = sc.params.copy()
    p0.mu = 1000.0 * nu0
    p0.gravity[0] = g0_factor * g_star
    return dict(scenario=sc, params_star=p_star, params0=p0, g_star=g_star, nu_star=nu_star, nu0=nu0,
                steps=steps, checkpoint_every=checkpoint_every, bins=bins)
