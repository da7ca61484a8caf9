"""System objects on the GPU against the oracle's update(): the reference
classes through their Python mirror, HIP-graph replay, and the state API."""
import numpy as np
import pytest

from cmag_amd import _lib as L
from cmag_amd import scenarios as S
from cmag_amd.systems import DamBreakSystem, ParamSystem, PeristalsisSystem, PoiseuilleFlowSystem

pytestmark = pytest.mark.gpu


def oracle_after(oracle, sysobj_params, pos, steps, vel=None, measures=None):
    sim = oracle.Simulation(sysobj_params, pos, vel, vel, measures)
    sim.update(steps)
    return sim


def check_grid_state(sysobj, sim):
    """hash, index, cellStart bit-exact; cellEnd on occupied cells"""
    assert np.array_equal(sysobj.getUintArray(L.ARRAY_HASH), sim.hash)
    assert np.array_equal(sysobj.getUintArray(L.ARRAY_INDEX), sim.index)
    cs = sysobj.getUintArray(L.ARRAY_CELLSTART)
    assert np.array_equal(cs, sim.cellStart)
    occ = cs != L.CELL_EMPTY
    assert np.array_equal(sysobj.getUintArray(L.ARRAY_CELLEND)[occ], sim.cellEnd[occ])


def rel_field(got, want, cols=3):
    scale = np.abs(want[:, :cols]).max()
    return float(np.abs(got[:, :cols].astype(np.float64) - want[:, :cols]).max() / max(scale, 1e-30))


def test_dambreak_update_matches_oracle(device, oracle):
    fluid, bo, grid, r = (64, 64, 1), 1, (256, 128, 4), 1 / 256
    s = DamBreakSystem(fluid, bo, grid, r, False)
    s.reset()
    p = s.params
    pos0 = s.getArray(DamBreakSystem.POSITION)
    assert np.array_equal(pos0, oracle.dambreak_init(p, s.getNumParticles()))
    s.update()
    sim = oracle_after(oracle, p, pos0, 1)
    check_grid_state(s, sim)
    assert np.allclose(s.getArray(DamBreakSystem.MEASURES)[:, 0], sim.measures[:, 0], rtol=1e-5)
    assert rel_field(s.getArray(DamBreakSystem.ACCELERATION), sim.acc) <= 1e-5
    assert rel_field(s.getArray(DamBreakSystem.VELOCITY), sim.vel) <= 1e-5
    for _ in range(19):
        s.update()
    sim.update(19)
    check_grid_state(s, sim)
    assert rel_field(s.getArray(DamBreakSystem.POSITION), sim.pos) <= 1e-6
    assert rel_field(s.getArray(DamBreakSystem.VELOCITY), sim.vel) <= 1e-3
    assert s.getElapsedTime() == pytest.approx(20 * 1e-4, rel=1e-5)


def test_poiseuille_report_run_matches_oracle(device, oracle):
    """The Poiseuille.Report configuration (dump.h:17-27), 225 steps (t = 0.0225)."""
    r = 1.0 / (2 * (64 - 2 * 3) * 1000)
    s = PoiseuilleFlowSystem((16, 58, 1), 3, (16, 64, 4), r, False)
    s.reset()
    p = s.params
    pos0 = s.getArray(0)
    sim = oracle_after(oracle, p, pos0, 0)
    steps = 0
    while s.getElapsedTime() < 0.0225:
        s.update()
        steps += 1
    sim.update(steps)
    check_grid_state(s, sim)
    v = s.getArray(PoiseuilleFlowSystem.VELOCITY)
    assert rel_field(v, sim.vel, cols=1) <= 1e-4
    assert rel_field(s.getArray(0), sim.pos) <= 1e-6


def test_peristalsis_update_matches_oracle(device, oracle):
    r = 1.0 / (2 * (64 - 6) * 1000)
    s = PeristalsisSystem(1e-4, (64, 58, 1), 0.6 * 35 * r, 100 * 1e-4, 1e-4, (0, 0, 0), 3, (64, 128, 4), r, False)
    s.Reset()
    p = s.params
    assert p.particle_mass > 0
    pos0 = s.getArray(PeristalsisSystem.POSITION)
    sim = oracle.Simulation(p, pos0, None, None, np.zeros_like(pos0))
    s.Update()
    sim.update(1)
    check_grid_state(s, sim)
    assert rel_field(s.getArray(PeristalsisSystem.VISCOUSFORCE), sim.acc) <= 1e-5
    assert rel_field(s.getArray(PeristalsisSystem.PRESSUREFORCE), sim.presF) <= 1e-5
    # trajectories: forces built from velocity differences amplify rounding as
    # steps accumulate, positions and velocities stay close
    for _ in range(29):
        s.Update()
    sim.update(29)
    check_grid_state(s, sim)
    assert rel_field(s.getArray(PeristalsisSystem.POSITION), sim.pos) <= 1e-6
    assert rel_field(s.getArray(PeristalsisSystem.VELOCITY), sim.vel) <= 1e-3
    assert s.GetElapsedTime() == pytest.approx(sim.elapsed, rel=1e-6)


@pytest.mark.parametrize("order", [L.HASH_LINEAR, L.HASH_MORTON])
def test_dam3d_config_system_matches_oracle(device, oracle, order):
    sc = S.dambreak3d(dx=0.02, hash_order=order)
    s = ParamSystem.from_state(sc.params, sc.pos, sc.vel)
    sim = oracle.Simulation(sc.params, sc.pos)
    s.update()
    sim.update(1)
    check_grid_state(s, sim)
    assert rel_field(s.getArray(L.ARRAY_ACCELERATION)[:sc.num_fluid], sim.acc[:sc.num_fluid]) <= 1e-5
    for _ in range(9):
        s.update()
    sim.update(9)
    check_grid_state(s, sim)
    assert rel_field(s.getArray(L.ARRAY_POSITION), sim.pos) <= 1e-6


@pytest.mark.parametrize("order,xbits", [(L.HASH_MORTON, 4), (L.HASH_LINEAR, 0), (L.HASH_MORTON, 0)])
def test_subcell_walk_matches_plain_walk_and_oracle(device, oracle, order, xbits):
    """the sub-cell walk order against the reference-order walk and the oracle:
    hash, index and cell ranges bit for bit, the sorted rows and the order of
    measures unchanged, density and force equal to rounding (other summation
    order), and the same trajectory over 10 steps (graph replay included)"""
    sc = S.dambreak3d(dx=0.02, hash_order=order, morton_xbits=xbits)
    a = ParamSystem.from_state(sc.params, sc.pos, sc.vel)
    b = ParamSystem.from_state(sc.params, sc.pos, sc.vel)
    assert a.subcell_walk
    b.set_subcell_walk(False)
    assert not b.subcell_walk
    sim = oracle.Simulation(sc.params, sc.pos)
    a.update()
    b.update()
    sim.update(1)
    check_grid_state(a, sim)
    for arr in (L.ARRAY_HASH, L.ARRAY_INDEX):
        assert np.array_equal(a.getUintArray(arr), b.getUintArray(arr))
    assert np.array_equal(a.getArray(L.ARRAY_SORTEDPOS), b.getArray(L.ARRAY_SORTEDPOS))
    ma, mb = a.getArray(L.ARRAY_MEASURES), b.getArray(L.ARRAY_MEASURES)
    assert np.max(np.abs(ma[:, 0] - mb[:, 0]) / mb[:, 0]) <= 1e-6
    assert np.max(np.abs(ma[:, 0] - sim.measures[:, 0]) / sim.measures[:, 0]) <= 1e-5
    nf = sc.num_fluid
    acc_a, acc_b = a.getArray(L.ARRAY_ACCELERATION)[:nf], b.getArray(L.ARRAY_ACCELERATION)[:nf]
    # the two walks sum the gate's Lennard-Jones terms (the field's largest)
    # in different orders
    assert rel_field(acc_a, acc_b) <= 3e-6
    assert rel_field(acc_a, sim.acc[:nf]) <= 1e-5
    a.run(5, use_graph=True)
    a.run(4, use_graph=False)
    b.run(9, use_graph=False)
    sim.update(9)
    check_grid_state(a, sim)
    assert rel_field(a.getArray(L.ARRAY_POSITION), b.getArray(L.ARRAY_POSITION)) <= 1e-6
    assert rel_field(a.getArray(L.ARRAY_POSITION), sim.pos) <= 1e-6


def test_subcell_walk_on_eighth_boundaries(device, oracle):
    """particles exactly on the eighth boundaries of their cells (x on a
    lattice of cell/8, the boundaries of the sub-cell walk's buckets), fluid and
    wall mixed in every cell: the sub-cell walk equals the reference-order walk
    to rounding and the oracle to 1e-5"""
    sc = S.dambreak3d(dx=0.02, hash_order=L.HASH_MORTON, morton_xbits=4)
    p = sc.params
    cs = float(p.cell_size[0])
    rng = np.random.default_rng(7)
    n = 6000
    x0 = float(p.world_origin[0]) + cs * np.floor((0.05 - float(p.world_origin[0])) / cs)
    pos = np.zeros((n, 4), np.float32)
    pos[:, 0] = x0 + (cs / 8) * rng.integers(0, 80, n)
    pos[:, 1] = rng.uniform(0.0, 0.3, n)
    pos[:, 2] = rng.uniform(0.0, 0.3, n)
    pos[:, 3] = np.where(rng.uniform(size=n) < 0.2, 1.0, 4.0)  # CMAG_DB_FIRST walls, fluid
    a = ParamSystem.from_state(p, pos, None)
    b = ParamSystem.from_state(p, pos, None)
    b.set_subcell_walk(False)
    sim = oracle.Simulation(p, pos)
    a.update()
    b.update()
    sim.update(1)
    check_grid_state(a, sim)
    ma, mb = a.getArray(L.ARRAY_MEASURES), b.getArray(L.ARRAY_MEASURES)
    assert np.max(np.abs(ma[:, 0] - mb[:, 0]) / mb[:, 0]) <= 1e-6
    assert np.max(np.abs(ma[:, 0] - sim.measures[:, 0]) / sim.measures[:, 0]) <= 1e-5
    fl = pos[:, 3] == 4.0
    acc_a, acc_b = a.getArray(L.ARRAY_ACCELERATION)[fl], b.getArray(L.ARRAY_ACCELERATION)[fl]
    assert rel_field(acc_a, acc_b) <= 1e-6
    assert rel_field(acc_a, sim.acc[fl]) <= 1e-5


@pytest.mark.parametrize("n", [1, 2, 3, 7, 65, 301, 4099])
def test_subcell_walk_small_and_ragged(device, oracle, n):
    """tiny and ragged particle counts (one partial block, runs of one row,
    single-particle cells) in a dam-break grid, fluid and walls mixed: the
    sub-cell walk equals the reference-order walk to rounding and the oracle
    to 1e-5"""
    sc = S.dambreak3d(dx=0.02, hash_order=L.HASH_MORTON, morton_xbits=4)
    p = sc.params
    rng = np.random.default_rng(n)
    pos = np.zeros((n, 4), np.float32)
    span = 0.06 if n < 100 else 0.3
    pos[:, :3] = rng.uniform(0.0, span, (n, 3))
    pos[:, 3] = np.where(rng.uniform(size=n) < 0.2, 1.0, 4.0)  # CMAG_DB_FIRST walls, fluid
    a = ParamSystem.from_state(p, pos, None)
    b = ParamSystem.from_state(p, pos, None)
    assert a.subcell_walk
    b.set_subcell_walk(False)
    sim = oracle.Simulation(p, pos)
    a.update()
    b.update()
    sim.update(1)
    check_grid_state(a, sim)
    ma, mb = a.getArray(L.ARRAY_MEASURES), b.getArray(L.ARRAY_MEASURES)
    assert np.max(np.abs(ma[:, 0] - mb[:, 0]) / mb[:, 0]) <= 1e-6
    assert np.max(np.abs(ma[:, 0] - sim.measures[:, 0]) / sim.measures[:, 0]) <= 1e-5
    fl = pos[:, 3] == 4.0
    if fl.any():
        acc_a, acc_b = a.getArray(L.ARRAY_ACCELERATION)[fl], b.getArray(L.ARRAY_ACCELERATION)[fl]
        assert rel_field(acc_a, acc_b) <= 1e-6
        assert rel_field(acc_a, sim.acc[fl]) <= 1e-5


def test_subcell_walk_independent_of_hash_order(device):
    """the sub-cell walk visits each target's candidates in the same order
    whatever the cell order (a stencil row's cells in x order, a cell's rows by
    bucket and then by particle), so its trajectory is the same bit for bit
    under the linear order, plain Z order and 16- and 128-cell bricks.  These
    orders split stencil rows into runs differently and stage the cell lookups
    of different blocks in LDS (a block is staged when its targets share one
    row of cells whose stencil range lies on consecutive keys); the other
    blocks look their runs up in global memory."""
    states = []
    for order, xbits in ((L.HASH_LINEAR, 0), (L.HASH_MORTON, 0), (L.HASH_MORTON, 4), (L.HASH_MORTON, 7)):
        sc = S.dambreak3d(dx=0.02, hash_order=order, morton_xbits=xbits)
        s = ParamSystem.from_state(sc.params, sc.pos, sc.vel)
        assert s.subcell_walk
        s.run(10, use_graph=False)
        states.append([s.getArray(a) for a in (L.ARRAY_POSITION, L.ARRAY_VELOCITY, L.ARRAY_VELOCITYLEAPFROG)])
    assert all(np.isfinite(x).all() for x in states[0])
    for st in states[1:]:
        for got, want in zip(st, states[0]):
            assert np.array_equal(got, want)


def _mixed_box(p, n, seed, span=0.3, eighth_lattice=False):
    rng = np.random.default_rng(seed)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(0.0, span, (n, 3))
    if eighth_lattice:
        cs = float(p.cell_size[0])
        x0 = float(p.world_origin[0]) + cs * np.floor((0.05 - float(p.world_origin[0])) / cs)
        pos[:, 0] = x0 + (cs / 8) * rng.integers(0, 80, n)
    pos[:, 3] = np.where(rng.uniform(size=n) < 0.2, 1.0, 4.0)  # CMAG_DB_FIRST walls, fluid
    return pos


@pytest.mark.parametrize("case", ["dam3d", "dam3d_linear", "eighth_boundaries", "ragged", "gated_1m", "dam2d"])
def test_inrange_masks_change_nothing(device, case):
    """The force pass fed by the density pass's in-range masks (wall-free rows
    of staged blocks) pushes the same neighbours in the same order as the
    force pass that tests every candidate: the trajectories are the same bit
    for bit, eager and graph-replayed, with walls mixed into cells, particles
    on eighth boundaries, ragged counts, the gated 1M dam break and a 2-D
    dam break."""
    if case == "dam2d":
        # the reference's 2-D dam break (64 x 64 column) on cells of 2h
        # (cellcount 1), so that its blocks stage their lookups and use masks
        import ctypes as C
        from oracle import oracle as O
        fluid, bo, grid, r = (64, 64, 1), 1, (256, 128, 4), 1.0 / 128
        p = S.params_dambreak(fluid, bo, grid, r)
        p.cellcount = 1
        for a in range(3):
            p.cell_size[a] = 2.0 * p.smoothing_radius
        L.call("cmag_params_finalize", C.byref(p))
        pos, vel, steps = O.dambreak_init(p, O.dambreak_count(fluid, bo, grid)), None, 200
    elif case == "gated_1m":
        sc = S.dambreak3d(dx=5e-3, hash_order=L.HASH_MORTON, morton_xbits=7)
        # long enough for rows whose mask walk fills a queue to be followed
        # by walked rows (the surge meets the floor and the far wall)
        p, pos, vel, steps = sc.params, sc.pos, sc.vel, 80
    if case not in ("dam2d", "gated_1m"):
        order, xbits = (L.HASH_LINEAR, 0) if case == "dam3d_linear" else (L.HASH_MORTON, 4)
        sc = S.dambreak3d(dx=0.02, hash_order=order, morton_xbits=xbits)
        p, vel, steps = sc.params, None, 6
        pos = {"dam3d": sc.pos, "dam3d_linear": sc.pos,
               "eighth_boundaries": _mixed_box(p, 6000, 7, eighth_lattice=True),
               "ragged": _mixed_box(p, 4099, 4099)}[case]
    print(f"{case}: {pos.shape[0]} particles", flush=True)
    a = ParamSystem.from_state(p, pos, vel)
    b = ParamSystem.from_state(p, pos, vel)
    assert a.subcell_walk and a.inrange_masks
    b.set_inrange_masks(False)
    assert not b.inrange_masks
    a.update()
    b.update()
    assert np.array_equal(a.getArray(L.ARRAY_ACCELERATION), b.getArray(L.ARRAY_ACCELERATION))
    a.run(steps, use_graph=True)
    b.run(steps, use_graph=False)
    # the random boxes put particles on top of each other; they blow up (NaN)
    # within a few steps, the same way in both
    for arr in (L.ARRAY_POSITION, L.ARRAY_VELOCITY, L.ARRAY_VELOCITYLEAPFROG, L.ARRAY_MEASURES):
        ga, gb = a.getArray(arr), b.getArray(arr)
        print(f"array {arr}: {int(np.sum(~np.isfinite(ga)))} / {int(np.sum(~np.isfinite(gb)))} non-finite, "
              f"{int(np.sum((ga != gb) & np.isfinite(ga) & np.isfinite(gb)))} finite entries differ", flush=True)
        assert np.array_equal(ga, gb, equal_nan=True), arr
    a.close()
    b.close()


def test_subcell_walk_rejected_where_it_does_not_apply(device):
    sc = S.poiseuille_channel()
    s = ParamSystem.from_state(sc.params, sc.pos, sc.vel)
    assert not s.subcell_walk
    with pytest.raises(RuntimeError, match="sub-cell"):
        s.set_subcell_walk(True)
    s.set_subcell_walk(False)


@pytest.mark.parametrize("case", ["channel", "dam3d_subcell"])
def test_graph_replay_equals_eager(device, case):
    sc = S.poiseuille_channel() if case == "channel" else S.dambreak3d(dx=0.02, hash_order=L.HASH_MORTON,
                                                                          morton_xbits=4)
    a = ParamSystem.from_state(sc.params, sc.pos, sc.vel)
    b = ParamSystem.from_state(sc.params, sc.pos, sc.vel)
    a.run(25, use_graph=False)
    b.run(25, use_graph=True)
    for arr in (L.ARRAY_POSITION, L.ARRAY_VELOCITY, L.ARRAY_VELOCITYLEAPFROG):
        assert np.array_equal(a.getArray(arr), b.getArray(arr))
    assert a.getElapsedTime() == b.getElapsedTime()
    # eager work between replays, then more replays
    b.run(5, use_graph=True)
    b.update()
    b.run(5, use_graph=True)
    a.run(11, use_graph=False)
    assert np.array_equal(a.getArray(L.ARRAY_POSITION), b.getArray(L.ARRAY_POSITION))


def test_peristalsis_graph_through_configuration(device, oracle):
    """a Peristalsis channel through its boundary-configuration phase (3-step
    wall ramp, 1000 dt of relaxation, switch to the moving wave) replayed from
    HIP graphs, with eager steps between replays, equals eager update() bit for
    bit; the phase switch and clock follow the oracle's state machine and the
    walls sit where the oracle puts them"""
    sc = S.channel_peristalsis(1500, 0.3, 1.0, seed=302)
    p = sc.params.copy()
    p.is_boundary_configuration = 1
    p.amplitude = np.float32(2.5 * p.delta_time * 1e-3)
    a = ParamSystem.from_state(p, sc.pos, None)
    b = ParamSystem.from_state(p, sc.pos, None)
    sim = oracle.Simulation(p, sc.pos)
    for chunk in (3, 997):
        a.run(chunk, use_graph=False)
        b.run(chunk, use_graph=True)
        sim.update(chunk)
        assert a.params.is_boundary_configuration == b.params.is_boundary_configuration == 1
    a.run(10, use_graph=False)
    b.run(4, use_graph=True)
    b.update()
    b.run(5, use_graph=True)
    sim.update(10)
    assert a.params.is_boundary_configuration == b.params.is_boundary_configuration == 0
    assert sim.P.is_boundary_configuration == 0
    for arr in (L.ARRAY_POSITION, L.ARRAY_VELOCITY, L.ARRAY_VELOCITYLEAPFROG):
        assert np.array_equal(a.getArray(arr), b.getArray(arr))
    assert a.getElapsedTime() == b.getElapsedTime() == np.float32(sim.elapsed)
    pos = b.getArray(L.ARRAY_POSITION)
    walls = sc.pos[:, 3] != 0.0
    assert np.abs(pos[walls, 1] - sim.pos[walls, 1]).max() <= 1e-6 * np.abs(sim.pos[walls, 1]).max()


def test_set_get_array_roundtrip_and_errors(device):
    s = DamBreakSystem((16, 16, 1), 1, (64, 32, 4), 1 / 64, False)
    s.reset()
    n = s.getNumParticles()
    data = np.arange(n * 4, dtype=np.float32).reshape(n, 4)
    s.setArray(DamBreakSystem.VELOCITY, data)
    assert np.array_equal(s.getArray(DamBreakSystem.VELOCITY), data)
    s.setArray(DamBreakSystem.VELOCITY, data[:5] * 2, start=10, count=5)
    assert np.array_equal(s.getArray(DamBreakSystem.VELOCITY, 10, 5), data[:5] * 2)
    with pytest.raises(L.CmagError):
        s.setArray(DamBreakSystem.VELOCITY, data, start=1)
    with pytest.raises(L.CmagError):
        s.getArray(99)
    with pytest.raises(L.CmagError):
        ParamSystem(sc_bad_params(), 10)


def sc_bad_params():
    p = S.poiseuille_channel().params.copy()
    p.grid_size[1] = 100
    return p


def test_right_boundary_controls_on_device(device, oracle):
    s = DamBreakSystem((32, 32, 1), 1, (128, 64, 4), 1 / 64, False)
    s.reset()
    p = s.params
    pos = s.getArray(0)
    s.changeRightBoundary()
    assert np.array_equal(s.getArray(0), oracle.shift_right_boundary(p, pos))
    assert s.params.right_boundary == pytest.approx(p.right_boundary + 32 * 2 / 64)
    s.update()
    s.removeRightBoundary()
    assert s.getElapsedTime() == 0.0
    assert s.params.right_boundary == np.float32(0xFFFFFFFF)


def test_channel_2000_steps_matches_series_and_fp64_oracle(device, oracle):
    """The CPU-runnable channel (32x64 fluid, Tait, Morris walls, h = 1.3 dx)
    run for 2000 CFL steps on the GPU (HIP graph): its 64-bin x-velocity profile
    is within 1% L2 of the analytic start-up series at the same time T (taken at
    K*nu, K = the lattice's Laplacian moment) and within 1e-3 of the FP64
    oracle's profile after the same 2000 steps."""
    sc = S.poiseuille_channel()
    steps = 2000
    s = ParamSystem.from_state(sc.params, sc.pos, sc.vel)
    s.run(steps, use_graph=True)
    pos, vel = s.getArray(L.ARRAY_POSITION), s.getArray(L.ARRAY_VELOCITY)

    def profile(p, v):
        fl = p[:, 3] == 0
        y = p[fl, 1].astype(np.float64) - sc.meta["y0"]
        b = np.clip((y / sc.meta["width"] * 64).astype(int), 0, 63)
        return np.bincount(b, v[fl, 0], 64) / np.maximum(np.bincount(b, None, 64), 1)

    got = profile(pos, vel)
    T = steps * float(np.float32(sc.params.delta_time))
    yc = (np.arange(64) + 0.5) / 64 * sc.meta["width"]
    k = S.lattice_laplacian_moment(sc.meta["h"], sc.meta["dx"], 2)
    want = S.poiseuille_series(yc, T, sc.meta["g"], sc.meta["width"], sc.meta["nu"] * k)
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    # against the series at the stated nu the Morris term's lattice moment
    # (K = 0.974 at h = 1.3 dx) shows: 1.6 % in the FP32 oracle's run
    raw = S.poiseuille_series(yc, T, sc.meta["g"], sc.meta["width"], sc.meta["nu"])
    err_raw = np.linalg.norm(got - raw) / np.linalg.norm(raw)
    print(f"channel L2 vs series: {err:.4f} at K*nu (K = {k:.4f}), {err_raw:.4f} at nu", flush=True)
    assert err < 0.01
    assert err_raw < 0.025
    sim = oracle.Simulation(sc.params, sc.pos, sc.vel, sc.vel, dtype=np.float64)
    sim.update(steps)
    ref = profile(sim.pos, sim.vel)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-3
