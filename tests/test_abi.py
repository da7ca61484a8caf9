"""Host-side surface of libcmag_hip.so: loads, exports every declared entry point,
parameter blocks match the reference constructors, reference initial states
match the oracle bit for bit.  No device work (CPU only)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from cmag_amd import _lib as L
from cmag_amd import scenarios as S

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(cmag_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_library_loads_and_exports_every_declared_symbol():
    lib = L.lib()
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding types every one of them
    assert not (set(names) - set(L.SIGNATURES)), set(names) - set(L.SIGNATURES)


def test_params_layout():
    assert L.lib().cmag_params_size() == C.sizeof(L.CmagParams)
    assert C.sizeof(L.CmagParams) % 4 == 0


def test_dambreak_params_follow_reference_ctor():
    """fluidSystem.cpp:40-75 for the demo configuration (128x128 fluid, bo 1,
    grid 512x256x4, r = 1/256)."""
    p = S.params_dambreak((128, 128, 1), 1, (512, 256, 4), 1 / 256)
    f32 = np.float32
    r = f32(1 / 256)
    assert p.smoothing_radius == f32(3.0) * r
    assert p.particle_mass == f32(1000.0 / 16327.6)
    assert p.cellcount == 3
    assert p.world_origin[0] == -(f32(512) * r) and p.world_origin[1] == -(f32(256) * r)
    assert p.cell_size[0] == r * f32(2.0)
    assert p.gravity[1] == f32(-9.8) and p.boundary_damping == -1.0
    B = f32(200) * f32(1000) * f32(9.8) * (f32(2) * r * f32(128)) / f32(7)
    assert p.B == pytest.approx(B, rel=1e-7)
    assert p.lj_D == pytest.approx(10 * -9.8 * 2 * (1 / 256) * 128, rel=1e-6)
    assert p.lj_a == r
    assert p.soundspeed == pytest.approx(np.sqrt(B * 7 / 1000), rel=1e-6)
    assert p.delta_time == f32(1e-4)
    assert p.right_boundary == pytest.approx(-2.0 + (1 + 128) * 2 * (1 / 256), rel=1e-7)
    assert p.clamp_lo[0] == pytest.approx(-2.0 + 2 / 256) and p.clamp_lo[1] == pytest.approx(-1.0 + 2 / 256)
    assert p.eos == L.EOS_TAIT and p.gridpos_rule == L.GRIDPOS_DIV_CELL and p.periodic_mask == 0
    assert p.kernel_coeff == pytest.approx(7 / (4 * np.pi * (3 / 256) ** 2), rel=1e-6)


def test_poiseuille_params_follow_reference_ctor():
    """poiseuilleFlowSystem.cpp:38-65 for the report configuration."""
    r = 1.0 / (2 * 58 * 1000)
    p = S.params_poiseuille((16, 58, 1), 3, (16, 64, 4), r)
    assert p.particle_mass == np.float32(1000.0 / 3381320880.551724)
    assert p.gravity[0] == np.float32(1e-4) and p.soundspeed == np.float32(1e-4)
    assert p.mu == np.float32(1e-3) and p.delta_time == np.float32(1e-4)
    assert p.periodic_mask == 1 and p.eos == L.EOS_LINEAR_C2
    assert p.gridpos_rule == L.GRIDPOS_HALF_DIV_RADIUS
    assert p.wall_top - p.wall_bottom == pytest.approx(58 * 2 * r, rel=1e-5)


def test_peristalsis_params_follow_reference_ctor():
    """PeristalsisSystem.cpp:26-57 with the demo arguments."""
    r = 1.0 / (2 * (64 - 6) * 1000)
    p = S.params_peristalsis(1e-4, (256, 58, 1), 0.6 * 35 * r, 100 * 1e-4, 1e-4, (0, 0, 0), 3,
                             (256, 128, 4), r)
    assert p.periodic_mask == 7 and p.eos == L.EOS_LINEAR_50C
    assert p.is_boundary_configuration == 1 and p.particle_mass == 0.0
    assert p.world_size[0] == pytest.approx(256 * 2 * r, rel=1e-6)
    assert p.wave_k == pytest.approx(np.pi / (256 * r), rel=1e-6)
    assert p.mu == np.float32(1e-3)


def test_invalid_params_rejected():
    p = S.params_dambreak((8, 8, 1), 1, (32, 16, 4), 0.5)
    p.grid_size[0] = 30
    with pytest.raises(L.CmagError):
        S.finalize(p)
    p = S.params_dambreak((8, 8, 1), 1, (32, 16, 4), 0.5)
    p.gamma = 7.5
    with pytest.raises(L.CmagError):
        S.finalize(p)


@pytest.mark.parametrize("kind", ["dambreak", "poiseuille", "peristalsis"])
def test_reference_initial_state_matches_oracle(oracle, kind):
    if kind == "dambreak":
        fluid, bo, grid, r = (40, 30, 1), 3, (128, 64, 4), 1 / 128
        p = S.params_dambreak(fluid, bo, grid, r)
        n = oracle.dambreak_count(fluid, bo, grid)
        want = oracle.dambreak_init(p, n)
    elif kind == "poiseuille":
        p = S.params_poiseuille((16, 58, 1), 3, (16, 64, 4), 1.0 / (2 * 58 * 1000))
        n = 16 * 58 + 2 * 16 * 3
        want = oracle.poiseuille_init(p, n)
    else:
        r = 1.0 / (2 * (64 - 6) * 1000)
        p = S.params_peristalsis(1e-4, (64, 58, 1), 0.6 * 35 * r, 1e-2, 1e-4, (0, 0, 0), 3, (64, 128, 4), r)
        n = 64 * 58 + 2 * 64 * 3
        want = oracle.peristalsis_init(p, n)
    got = np.zeros((n, 4), np.float32)
    L.call("cmag_reference_initial_state", C.byref(p), n, got.ctypes.data)
    assert np.array_equal(got, want)
    if kind == "peristalsis":
        s = L.lib().cmag_peristalsis_lattice_sum(C.byref(p), got.ctypes.data, n)
        assert s == oracle.peristalsis_mass_sum(p, want)


def test_scenarios_geometry():
    sc = S.poiseuille_channel()
    assert sc.pos.shape == (2048 + 192, 4) and sc.num_fluid == 2048
    p = sc.params
    assert tuple(p.grid_size) == (32, 128, 4) and p.cellcount == 3
    assert p.delta_time == pytest.approx(5.157e-3, rel=1e-3)
    fl = sc.pos[:2048]
    assert np.max(np.abs(fl[:, 0] - np.round((fl[:, 0] - p.world_origin[0]) * 64 - 0.5) / 64
                         - p.world_origin[0] - 0.5 / 64)) <= 1.01e-3 / 64
    db = S.dambreak3d(dx=0.02)
    assert db.num_fluid == 20 * 40 * 20
    assert np.all(db.pos[:db.num_fluid, 3] == 4.0)
    w = db.pos[db.num_fluid:]
    assert np.all((w[:, 3] == 1.0) | (w[:, 3] == 0.0))
    # the gate (RightFirst walls) fills the tank's cross-section at the column's face,
    # three layers thick, and the release shifts put it on the tank's far wall
    gate = w[w[:, 3] == 0.0]
    assert gate.shape[0] == 3 * 50 * 20 and db.meta["release_shifts"] == 3
    assert np.allclose(np.unique(gate[:, 0]), 0.4 + (np.arange(3) + 0.5) * 0.02)
    assert db.params.right_boundary == np.float32(0.4)
    moved = gate[:, :3] + np.array([3 * 0.4, 0.0, 0.0], np.float32)
    open_tank = S.dambreak3d(dx=0.02, gate=False)
    fixed = open_tank.pos[open_tank.num_fluid:, :3]
    key = lambda a: set(map(tuple, np.round(a / 0.001).astype(np.int64)))  # noqa: E731
    assert key(moved) <= key(fixed) and len(key(np.concatenate([w[w[:, 3] == 1.0, :3], moved]))) == fixed.shape[0]
    assert np.unique(np.round(db.pos[:, :3] / 0.001).astype(np.int64), axis=0).shape[0] == db.pos.shape[0]
    big = S.dambreak3d(dx=5e-3)
    assert tuple(big.params.grid_size) == (128, 128, 64) and big.num_fluid == 1_024_000
    assert 0.5e6 < big.meta["num_wall"] < 0.6e6


def test_splitmix64_reference_value():
    """First output for seed 0 of splitmix64 (Steele, Lea, Flood 2014)."""
    g = S.SplitMix64(0)
    assert g.next_u64() == 0xE220A8397B1DCDAF
    u = S.SplitMix64(1973).uniform(1000)
    assert np.all((u >= 0) & (u < 1)) and abs(u.mean() - 0.5) < 0.05


def test_channel_batch_workload_geometry():
    """the batched workload: 256 channels, counts log-uniform in [1k, 64k],
    half Poiseuille / half peristaltic, each grid spanning its x-period exactly"""
    import math

    import numpy as np

    from cmag_amd import scenarios as S

    b = S.channel_batch()
    assert len(b) == 256
    kinds = [s.meta["kind"] for s in b]
    assert kinds[:128] == ["poiseuille"] * 128 and kinds[128:] == ["peristalsis"] * 128
    counts = np.array([s.pos.shape[0] for s in b])
    assert counts.min() >= 900 and counts.max() <= 70000
    assert 3e6 < counts.sum() < 5e6
    nus = np.array([s.meta["nu"] for s in b[:128]])
    assert nus.min() >= 1e-3 and nus.max() <= 1e-1
    for s in b:
        p = s.params
        dx = s.meta["dx"]
        period = s.meta["nx"] * dx
        assert abs(p.grid_size[0] * p.cell_size[0] - period) < 1e-6 * period
        # cell count from the exact (double) h and cell, as the reference sizes
        # h = 3r with cell = 2r: cellcount 3
        assert p.cellcount == math.ceil(2 * s.meta["h"] / (period / p.grid_size[0]) - 1e-9)
        assert p.grid_size[2] == 1 and p.dim == 2 and p.periodic_mask == 1
        inside = (s.pos[:, 1] - p.world_origin[1]) / p.cell_size[1]
        assert inside.min() > 0 and inside.max() < p.grid_size[1]
    for s in b[128:]:
        p = s.params
        assert 0.1 <= s.meta["amp_ratio"] <= 0.5 and 0.5 <= s.meta["wl_ratio"] <= 2.0
        # fluid strictly between the innermost wall layers at t = 0
        fl = s.pos[:, 3] == 0
        x, y = s.pos[fl, 0].astype(np.float64), s.pos[fl, 1].astype(np.float64)
        k = 2 * math.pi / s.meta["wavelength"]
        sn = np.sin(k * (x - float(p.world_origin[0])))
        H, A = p.fluid_size[1] * p.particle_radius, float(p.amplitude)
        assert np.all(y > -H + A * sn) and np.all(y < H - A * sn)
        assert np.all(np.isin(s.pos[~fl, 3], [2, 4, 6, -2, -4, -6]))
