"""Pins of the oracle's DamBreak force terms and Tait equation of state by
mathematics, and proof that each pin has teeth.  CPU only.

The DamBreak force (fluid_kernel.cu:172-292) is
    a_i = -sum_j m (p_i/rho_i^2 + p_j/rho_j^2 + Pi_ij) grad_i W_ij
with the Monaghan viscosity (fluid_kernel.cu:225-237)
    Pi_ij = -nu (v_ij . r_ij) / (r^2 + 0.001 h^2)  when v_ij . r_ij < 0, else 0,
    nu = 2 alpha h c / (rho_i + rho_j)
and the Tait equation of state (fluid_kernel.cu:168)
    p = B ((rho/rho0)^gamma - 1).
Checks used here, none of which re-evaluates the oracle's expressions:
  * SPH consistency in 3-D and 2-D: on a lattice at rho0 a linear pressure
    field p = p0 + b x gives the interior acceleration -b/rho0 times the
    lattice moment sum_j dx_j^2 |W'|/r dV, whose continuum value is 1 (the
    moment is within 0.1 % of 1 at h = 2 dx and within 3 % at h = 1.3 dx).
    This fixes the gradient normalisations 21/(8 pi h^4) (3-D) and
    7/(2 pi h^3) (2-D) and the sign of the pressure term.
  * two-particle closed forms with exact inputs: an approaching pair at
    q = 1 and at q = 0.1 (where the 0.001 h^2 regulariser is 10 % of r^2)
    against hand-evaluated values; a separating pair gives exactly 0.
  * Tait closed forms: p(rho0) = 0, p(2^(1/7) rho0) = B, p(2 rho0) = 127 B,
    dp/drho(rho0) = B gamma / rho0 by central differences, with densities set
    exactly through an isolated particle's own kernel value.
Each planted mistake (ORACLE_MUTATE in oracle/oracle_stages.inc: a factor of
2 in either gradient normalisation or in nu, 0.01 h^2 for the regulariser, the
viscosity branch on separating pairs, gamma - 1 as the exponent, 2B, a factor
2 in the 3-D kernel normalisation) must fail at least one pin.
"""
import math
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from cmag_amd import _lib as L
from test_oracle import grid_params, lattice_box

DB_FLUID = 4.0  # fluid type code, fluid_kernel.cuh:18-22


# ---------------------------------------------------------------- linear pressure field
def _linear_pressure(O, dim, hf, rtol_lo, rtol_hi):
    """interior a = -b/rho0 (times the lattice moment) for p = p0 + b x"""
    n, dx = (14, 1.0 / 16) if dim == 3 else (24, 1.0 / 32)
    cc = int(math.ceil(2 * hf))
    grid = (32, 32, 32) if dim == 3 else (32, 32, 1)
    p = grid_params(grid=grid, cs=dx, h=hf * dx, dim=dim, origin=(0, 0, 0), cellcount=cc)
    rho0 = 1000.0
    p.particle_mass = rho0 * dx ** dim
    p.soundspeed = 7.0
    pos = lattice_box(n, dx, dim)
    if dim == 2:
        pos[:, 2] = 0.5 * dx
    pos[:, 3] = DB_FLUID
    r = O.run_pipeline(p, pos, np.zeros_like(pos), dtype=np.float64)
    sp = r["sorted_pos"]
    meas = np.zeros((sp.shape[0], 4))
    meas[:, 0] = rho0
    b, p0 = 3.0e4, 5.0e3
    meas[:, 1] = p0 + b * sp[:, 0]
    a = O.force(p, meas, sp, r["sorted_vel"], r["index"], r["cell_start"], r["cell_end"], dtype=np.float64)
    k = np.rint(pos[:, :dim] / dx - 0.5).astype(int)   # lattice indices of the unsorted rows
    reach = int(math.ceil(2 * hf)) - 1                  # neighbours lie at most this many sites away
    interior = np.all((k >= reach) & (k <= n - 1 - reach), axis=1)
    assert interior.sum() >= 27 if dim == 3 else interior.sum() >= 9
    want = -b / rho0
    ratio = a[interior, 0] / want
    assert np.all(ratio > 1.0 - rtol_lo) and np.all(ratio < 1.0 + rtol_hi), (ratio.min(), ratio.max())
    for ax in range(1, dim):
        assert np.all(np.abs(a[interior, ax]) < 1e-6 * abs(want))


def pin_linear_pressure_3d(O):
    _linear_pressure(O, 3, 2.0, 2e-3, 2e-3)
    _linear_pressure(O, 3, 1.3, 0.03, 1e-3)


def pin_linear_pressure_2d(O):
    _linear_pressure(O, 2, 2.0, 2e-3, 2e-3)
    _linear_pressure(O, 2, 1.3, 0.03, 1e-3)


def test_dambreak_linear_pressure_field_3d(oracle):
    pin_linear_pressure_3d(oracle)


def test_dambreak_linear_pressure_field_2d(oracle):
    pin_linear_pressure_2d(oracle)


# ---------------------------------------------------------------- Monaghan viscosity
def _pair(O, dim, r, vi, vj, rho=(1000.0, 1048.0), alpha=0.5, c=4.0, h=0.5):
    """accelerations of two fluid particles at zero pressure, i at the origin
    side and j at +r along x (the unsorted order i, j)"""
    grid = (8, 8, 8) if dim == 3 else (8, 8, 1)
    p = grid_params(grid=grid, cs=2 * h, h=h, dim=dim, origin=(0, 0, 0), cellcount=1)
    p.particle_mass = 1.0
    p.alpha = alpha
    p.soundspeed = c
    pos = np.array([[2.0, 2.0, 2.0 if dim == 3 else 0.5, DB_FLUID],
                    [2.0 + r, 2.0, 2.0 if dim == 3 else 0.5, DB_FLUID]])
    vel = np.zeros_like(pos)
    vel[0, 0], vel[1, 0] = vi, vj
    res = O.run_pipeline(p, pos, vel, dtype=np.float64)
    meas = np.zeros((2, 4))
    meas[:, 0] = np.asarray(rho)[res["index"]]
    a = O.force(p, meas, res["sorted_pos"], res["sorted_vel"], res["index"], res["cell_start"], res["cell_end"],
                dtype=np.float64)
    return a


def pin_monaghan_3d(O):
    # h = 1/2, r = 1/2 (q = 1): nu = 2 * 0.5 * 0.5 * 4 / 2048 = 1/1024;
    # v_ij . r_ij = (1 - (-1)) * (-1/2) = -1, so Pi = (1/1024) / (1/4 + 1/4000);
    # grad W at q = 1: 21/(8 pi h^4) * ((1/2)^4 - (1/2)^3 * 3) = (42/pi) * (-5/16)
    # along r_ij/r = -x, so a_i,x = -m Pi (-1) (42/pi)(-5/16) = -Pi * 105/(8 pi)
    a = _pair(O, 3, 0.5, 1.0, -1.0)
    pi_ij = (1.0 / 1024.0) / (0.25 + 0.25 / 1000.0)
    want = -pi_ij * 105.0 / (8.0 * math.pi)
    assert a[0, 0] == pytest.approx(want, rel=1e-12)
    assert a[1, 0] == pytest.approx(-want, rel=1e-12)
    assert np.all(a[:, 1:3] == 0.0)


def pin_monaghan_2d(O):
    # the same pair in 2-D: grad W = 7/(2 pi h^3) * (-5/16) = (28/pi)(-5/16)
    a = _pair(O, 2, 0.5, 1.0, -1.0)
    pi_ij = (1.0 / 1024.0) / (0.25 + 0.25 / 1000.0)
    want = -pi_ij * 35.0 / (4.0 * math.pi)
    assert a[0, 0] == pytest.approx(want, rel=1e-12)
    assert a[1, 0] == pytest.approx(-want, rel=1e-12)


def pin_monaghan_regulariser(O):
    # r = 1/20 (q = 1/10), where 0.001 h^2 = 1/4000 is a tenth of r^2 = 1/400:
    # Pi = (1/1024) * (2/20) / (1/400 + 1/4000) = (1/1024) * (1/10) / (11/4000);
    # grad W: (42/pi) * (t^4 - t^3 * 6/5), t = 19/20
    a = _pair(O, 3, 0.05, 1.0, -1.0)
    pi_ij = (1.0 / 1024.0) * 0.1 / (11.0 / 4000.0)
    t = 19.0 / 20.0
    want = -pi_ij * (42.0 / math.pi) * -(t ** 4 - t ** 3 * 1.2)
    assert a[0, 0] == pytest.approx(want, rel=1e-12)


def pin_monaghan_separating(O):
    # a separating pair at zero pressure feels nothing, in both dimensions
    for dim in (3, 2):
        a = _pair(O, dim, 0.5, -1.0, 1.0)
        assert np.all(a[:, :3] == 0.0)


def test_monaghan_pair_closed_form(oracle):
    pin_monaghan_3d(oracle)
    pin_monaghan_2d(oracle)
    pin_monaghan_regulariser(oracle)
    pin_monaghan_separating(oracle)


def test_monaghan_viscosity_dissipates(oracle):
    """At zero pressure the viscous pair forces conserve momentum and never add
    kinetic energy: sum_i m a_i = 0 and sum_i m v_i . a_i < 0."""
    rng = np.random.default_rng(12)
    p = grid_params(grid=(16, 16, 16), cs=0.125, h=0.05, dim=3, origin=(0, 0, 0), cellcount=2)
    p.soundspeed = 10.0
    pos = lattice_box(8, 0.04, 3, origin=0.5)
    pos[:, :3] += rng.uniform(-0.004, 0.004, (pos.shape[0], 3))
    pos[:, 3] = DB_FLUID
    vel = np.zeros_like(pos)
    vel[:, :3] = rng.normal(size=(pos.shape[0], 3))
    r = oracle.run_pipeline(p, pos, vel, dtype=np.float64)
    meas = r["measures"].copy()
    meas[:, 1] = 0.0
    a = oracle.force(p, meas, r["sorted_pos"], r["sorted_vel"], r["index"], r["cell_start"], r["cell_end"],
                     dtype=np.float64)
    assert np.all(np.abs(a[:, :3].sum(0)) <= 1e-10 * np.abs(a[:, :3]).sum(0).max())
    power = float((vel[:, :3] * a[:, :3]).sum())
    assert power < 0.0


# ---------------------------------------------------------------- Tait
def _isolated_states(O, rhos, B=3.0, gamma=7.0, rho0=1000.0):
    """(rho, p) of an isolated particle whose density is set through the mass:
    one run at m = 1 gives the particle's own kernel value W0, then
    m = rho / W0.  The mass is an FP32 parameter, so rho lands within ~1e-7
    relative of the target; the checks below allow for that."""
    p = grid_params(grid=(8, 8, 8), cs=1.0, h=0.5, dim=3, origin=(0, 0, 0), cellcount=1)
    p.B, p.gamma, p.rest_density = B, gamma, rho0
    pos = np.array([[3.5, 3.5, 3.5, DB_FLUID]])
    p.particle_mass = 1.0
    w0 = float(O.run_pipeline(p, pos, np.zeros_like(pos), dtype=np.float64)["measures"][0, 0])
    out = []
    for rho in rhos:
        p.particle_mass = rho / w0
        m = O.run_pipeline(p, pos, np.zeros_like(pos), dtype=np.float64)["measures"][0]
        assert m[0] == pytest.approx(rho, rel=2e-7)
        out.append((float(m[0]), float(m[1])))
    return out


def pin_tait(O):
    B, rho0 = 3.0, 1000.0
    d = 1e-4
    (r0, p0), (rb, pb), (r2, p2), (rh, ph), (rl, pl) = _isolated_states(
        O, [rho0, 2.0 ** (1.0 / 7.0) * rho0, 2.0 * rho0, rho0 * (1 + d), rho0 * (1 - d)], B=B)
    # p(rho0) = 0: what is left is the slope times rho's own offset from rho0
    assert abs(p0) <= 7.0 * B * 4e-7
    assert pb == pytest.approx(B, rel=5e-6)          # (2^(1/7))^7 - 1 = 1
    assert p2 == pytest.approx(127.0 * B, rel=5e-6)  # 2^7 - 1 = 127
    # dp/drho at rho0 = B gamma / rho0 (central difference over the actual densities)
    assert (ph - pl) / (rh - rl) == pytest.approx(B * 7.0 / rho0, rel=2e-6)


def test_tait_closed_forms(oracle):
    pin_tait(oracle)


def test_tait_fp32_at_rest_density(oracle):
    """the FP32 restatement: p(rho0) is zero up to the rounding of rho0 * 1"""
    p = grid_params(grid=(8, 8, 8), cs=1.0, h=0.5, dim=3, origin=(0, 0, 0), cellcount=1)
    p.B, p.gamma, p.rest_density = 2.0e5, 7.0, 1000.0
    pos = np.array([[3.5, 3.5, 3.5, DB_FLUID]])
    p.particle_mass = 1.0
    w0 = float(oracle.run_pipeline(p, pos, np.zeros_like(pos), dtype=np.float32)["measures"][0, 0])
    p.particle_mass = 1000.0 / w0
    m = oracle.run_pipeline(p, pos, np.zeros_like(pos), dtype=np.float32)["measures"][0]
    assert abs(m[0] / 1000.0 - 1.0) < 4e-7
    assert abs(m[1]) < 7 * 4e-7 * 2.0e5 * 1.01


# ---------------------------------------------------------------- the pins have teeth
def _density_normalisation(O):
    # an isolated 3-D particle's density is m * 21/(16 pi h^3) (W integrates
    # to one; the 3-D normalisation is pinned independently by a fine-lattice
    # integral in test_oracle.py)
    p = grid_params(grid=(8, 8, 8), cs=1.0, h=0.5, dim=3, origin=(0, 0, 0), cellcount=1)
    p.particle_mass = 1.0
    pos = np.array([[3.5, 3.5, 3.5, DB_FLUID]])
    rho = O.run_pipeline(p, pos, np.zeros_like(pos), dtype=np.float64)["measures"][0, 0]
    h = 0.5
    dv = (2 * h / 40) ** 3
    g = (np.arange(-40, 40) + 0.5) * (2 * h / 40)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    q = np.sqrt(X ** 2 + Y ** 2 + Z ** 2) / h
    shape = np.where(q < 2, (1 - q / 2) ** 4 * (2 * q + 1), 0.0)
    # W(0) / integral of the unnormalised shape = the normalisation constant
    assert rho == pytest.approx(1.0 / (shape.sum() * dv), rel=2e-3)


PLANTED = {
    1: [pin_linear_pressure_3d],        # 3-D gradient normalisation 21/(16 pi h^4)
    2: [pin_linear_pressure_2d, pin_monaghan_2d],  # 2-D gradient 7/(4 pi h^3)
    3: [pin_monaghan_regulariser],      # regulariser 0.01 h^2
    4: [pin_monaghan_3d, pin_monaghan_2d],          # nu with a factor 1 instead of 2
    5: [pin_tait],                      # exponent gamma - 1
    6: [pin_tait],                      # 2 B
    7: [_density_normalisation],        # 3-D kernel 21/(8 pi h^3)
    8: [pin_monaghan_separating],       # viscosity on separating pairs
}


@pytest.fixture(scope="module")
def planted_builds(oracle, tmp_path_factory):
    d = tmp_path_factory.mktemp("planted")
    with ThreadPoolExecutor(max_workers=4) as ex:
        paths = list(ex.map(lambda k: oracle.build_variant(k, d), sorted(PLANTED)))
    return dict(zip(sorted(PLANTED), paths))


def test_density_normalisation_pin(oracle):
    _density_normalisation(oracle)


@pytest.mark.parametrize("k", sorted(PLANTED))
def test_pins_reject_planted_mistake(oracle, planted_builds, k):
    with oracle.using(planted_builds[k]):
        for pin in PLANTED[k]:
            with pytest.raises(AssertionError):
                pin(oracle)
