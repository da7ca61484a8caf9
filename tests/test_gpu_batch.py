"""Batched stepping on the GPU: every instance of a batch must step exactly as a
single System built from the same parameter block and state (bit for bit), on
the eager and the graph path, across the Peristalsis boundary-configuration
switch, and the autotuner must leave the state as it found it."""
import numpy as np
import pytest

from cmag_amd import _lib as L
from cmag_amd import scenarios as S
from cmag_amd.batch import Batch
from cmag_amd.systems import ParamSystem

pytestmark = pytest.mark.gpu

ARRAYS = (L.ARRAY_POSITION, L.ARRAY_VELOCITY, L.ARRAY_VELOCITYLEAPFROG)


def mixed_instances():
    """three kinds, uneven sizes; one Peristalsis channel still in its
    boundary-configuration phase, with a ramp short enough to finish"""
    rng = np.random.default_rng(5)
    out = []
    for k, n in enumerate((1500, 4200, 2600)):
        sc = S.channel_poiseuille(n, nu=[2e-3, 2e-2, 8e-2][k], seed=100 + k)
        vel = np.zeros_like(sc.pos)
        vel[:sc.num_fluid, :2] = rng.normal(scale=5e-3, size=(sc.num_fluid, 2))
        out.append((sc.params, sc.pos, vel))
    for k, (n, a, wl) in enumerate(((3300, 0.2, 0.7), (1800, 0.45, 1.8))):
        sc = S.channel_peristalsis(n, a, wl, seed=200 + k)
        out.append((sc.params, sc.pos, None))
    sc = S.channel_peristalsis(2000, 0.3, 1.0, seed=300)
    p = sc.params.copy()
    p.is_boundary_configuration = 1
    p.amplitude = np.float32(3.5 * p.delta_time * 1e-3)  # ramp of 4 steps, then 1000 dt of relaxation
    out.append((p, sc.pos, None))
    # a reference DamBreak block (2-D, linear hash) shares the batch
    p = S.params_dambreak((24, 20, 1), 3, (64, 32, 4), 1 / 128)
    from oracle import oracle as O
    n = O.dambreak_count((24, 20, 1), 3, (64, 32, 4))
    out.append((p, O.dambreak_init(p, n), None))
    return out


def singles(instances):
    """single Systems walking in reference order (the batch's order): a DamBreak
    System would otherwise take the sub-cell walk, whose sums agree only to
    rounding"""
    out = []
    for p, pos, vel in instances:
        s = ParamSystem.from_state(p, pos, vel)
        if s.subcell_walk:
            s.set_subcell_walk(False)
        out.append(s)
    return out


def assert_same(b, sys_list):
    for k, s in enumerate(sys_list):
        for a in ARRAYS:
            got, want = b.getArray(k, a), s.getArray(a)
            assert np.array_equal(got, want), (k, a, np.abs(got - want).max())
        assert b.elapsed(k) == s.getElapsedTime()


def test_batch_equals_single_systems_eager(device):
    inst = mixed_instances()
    b = Batch(inst)
    ss = singles(inst)
    for _ in range(12):
        b.update()
        for s in ss:
            s.update()
    assert_same(b, ss)


def test_batch_graph_and_launch_shapes(device):
    inst = mixed_instances()
    b = Batch(inst)
    ss = singles(inst)
    b.set_tuning(block=128, nb_cap=8, blocks_per_cu=2, steps_per_graph=4)
    b.run(10, use_graph=True)          # two graph launches + two eager steps
    for s in ss:
        for _ in range(10):
            s.update()
    assert_same(b, ss)
    b.set_tuning(block=64, nb_cap=32, blocks_per_cu=0, steps_per_graph=3)
    b.run(3, use_graph=True)
    b.update()                          # eager work between replays
    b.run(6, use_graph=True)
    for s in ss:
        for _ in range(10):
            s.update()
    assert_same(b, ss)


def test_configuration_switch_matches_single_system(device):
    """a Peristalsis channel through its whole boundary-configuration phase
    (ramp, 1000 dt of relaxation, switch to the moving wave)"""
    sc = S.channel_peristalsis(1500, 0.3, 1.0, seed=301)
    p = sc.params.copy()
    p.is_boundary_configuration = 1
    p.amplitude = np.float32(2.5 * p.delta_time * 1e-3)
    b = Batch([(p, sc.pos, None)])
    s = ParamSystem.from_state(p, sc.pos, None)
    b.run(1010, use_graph=True)
    for _ in range(1010):
        s.update()
    assert_same(b, [s])


def test_autotune_restores_state(device):
    inst = mixed_instances()
    b = Batch(inst)
    b.run(3, use_graph=False)
    before = [[b.getArray(k, a) for a in ARRAYS] for k in range(b.m)]
    t_before = [b.elapsed(k) for k in range(b.m)]
    tune, ms = b.autotune(steps_per_trial=2)
    assert tune["block"] in (64, 128, 256) and tune["nb_cap"] in (8, 16, 32) and ms > 0
    assert b.tuning == tune
    for k in range(b.m):
        for a, want in zip(ARRAYS, before[k]):
            assert np.array_equal(b.getArray(k, a), want)
        assert b.elapsed(k) == t_before[k]
    # and stepping continues as a fresh batch would
    ref = Batch(inst)
    ref.run(3, use_graph=False)
    b.run(5)
    ref.run(5, use_graph=False)
    for k in range(b.m):
        assert np.array_equal(b.getArray(k, L.ARRAY_POSITION), ref.getArray(k, L.ARRAY_POSITION))


def test_batch_of_3d_dam_breaks(device):
    """three 3-D dam breaks of different resolution in one batch: bit for bit
    the single Systems in reference-order walk, and to rounding the default
    (sub-cell walk) Systems, over 10 graph-replayed steps"""
    inst = []
    for dx in (0.05, 0.04, 0.06):
        sc = S.dambreak3d(dx=dx)
        inst.append((sc.params, sc.pos, sc.vel))
    b = Batch(inst)
    plain = singles(inst)
    sub = [ParamSystem.from_state(p, pos, vel) for p, pos, vel in inst]
    assert all(s.subcell_walk for s in sub)
    b.run(10, use_graph=True)
    for s in plain + sub:
        s.run(10, use_graph=False)
    assert_same(b, plain)
    for k, s in enumerate(sub):
        got, want = b.getArray(k, L.ARRAY_POSITION), s.getArray(L.ARRAY_POSITION)
        assert np.abs(got[:, :3] - want[:, :3]).max() <= 1e-6 * np.abs(want[:, :3]).max()


def test_batch_errors(device):
    sc2 = S.channel_poiseuille(1200, nu=1e-2, seed=1)
    sc3 = S.dambreak3d(dx=0.05)
    with pytest.raises(L.CmagError, match="dimension"):
        Batch([(sc2.params, sc2.pos, None), (sc3.params, sc3.pos, None)])
    b = Batch([(sc2.params, sc2.pos, None)])
    with pytest.raises(L.CmagError, match="count"):
        b.setArray(0, L.ARRAY_POSITION, sc2.pos[:-1])
    with pytest.raises(L.CmagError, match="instance"):
        L.call("cmag_batch_get_array", b._h, 5, L.ARRAY_POSITION, None, 0)
    with pytest.raises(L.CmagError, match="block"):
        b.set_tuning(block=96)


def test_channel_batch_256_full_workload(device, oracle):
    """The benchmarked batch itself: 256 channels (128 Poiseuille, 128
    peristaltic, 4.13M particles, so the global sort runs its large-tile path)
    at the benchmark's tuned launch shape.  Every instance equals its single
    System bit for bit after 20 graph-replayed steps, and after one step eight
    instances (both kinds, the smallest and the largest) equal the FP32 oracle's
    update(): densities to 1e-5 relative and positions to 1e-6.  Forces and
    velocities one step from rest are nearly cancelling pressure terms (a
    channel's jittered lattice), whose FP32 sums carry rounding of ~1e-4 of
    the net field in any summation order: there the GPU must be as close to
    the FP64 oracle as the FP32 oracle itself is (within 4x, or 1e-5)."""
    from tests.gpu_util import normwise
    scen = S.channel_batch()
    inst = [(sc.params, sc.pos, sc.vel) for sc in scen]
    sizes = np.array([sc.pos.shape[0] for sc in scen])
    print(f"channel batch: {len(scen)} instances, {sizes.sum():,} particles", flush=True)
    assert len(scen) == 256 and sizes.sum() > (1 << 18)
    b = Batch(inst)
    b.set_tuning(block=256, nb_cap=8, blocks_per_cu=0, steps_per_graph=10)
    # one eager step against the oracle on a sample of instances
    b.update()
    po, pe = np.arange(128), np.arange(128, 256)
    pick = {int(po[np.argmin(sizes[po])]), int(po[np.argmax(sizes[po])]),
            int(pe[np.argmin(sizes[pe])]), int(pe[np.argmax(sizes[pe])])}
    rng = np.random.default_rng(4)
    pick |= {int(k) for k in rng.choice(po, 2, replace=False)} | {int(k) for k in rng.choice(pe, 2, replace=False)}
    for k in sorted(pick):
        p, pos, vel = inst[k]
        sim = oracle.Simulation(p, pos, vel, vel)
        sim.update(1)
        s64 = oracle.Simulation(p, pos, vel, vel, dtype=np.float64)
        s64.update(1)
        gp, gv = b.getArray(k, L.ARRAY_POSITION), b.getArray(k, L.ARRAY_VELOCITY)
        gm, ga = b.getArray(k, L.ARRAY_MEASURES), b.getArray(k, L.ARRAY_ACCELERATION)
        fields = [(ga, sim.acc, s64.acc), (gv, sim.vel, s64.vel)]
        if k >= 128:
            fields.append((b.getArray(k, L.ARRAY_PRESSUREFORCE), sim.presF, s64.presF))
        ep = normwise(gp[:, :3], sim.pos[:, :3])
        ed = float(np.max(np.abs(gm[:, 0].astype(np.float64) - sim.measures[:, 0]) / sim.measures[:, 0]))
        msg = []
        for got, w32, w64 in fields:
            e_gpu, e_32 = normwise(got[:, :3], w64[:, :3]), normwise(w32[:, :3], w64[:, :3])
            msg.append(f"{e_gpu:.1e} (FP32 oracle {e_32:.1e})")
            assert e_gpu <= max(1e-5, 4.0 * e_32), (k, e_gpu, e_32)
        print(f"instance {k} ({sizes[k]} particles, {'Poiseuille' if k < 128 else 'peristaltic'}): "
              f"density {ed:.2e}, pos {ep:.2e}; vs FP64: force/vel/pressure {', '.join(msg)}", flush=True)
        assert ed <= 1e-5 and ep <= 1e-6
    # 20 more steps in two graphs of 10, against the single Systems
    ss = singles(inst)
    b.run(20, use_graph=True)
    for s in ss:
        for _ in range(21):
            s.update()
    assert_same(b, ss)
    for s in ss:
        s.close()
    b.close()
