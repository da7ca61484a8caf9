"""ctypes wrapper of oracle/liboracle.so.  TEST INFRASTRUCTURE ONLY.

The oracle is the checker: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may use it, never the product path.  See cmag_oracle.c for
what each restated function follows in the reference and how it is pinned.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_lib = None


def build() -> Path:
    """liboracle.so (AVX2, runs on any x86-64-v3 host) and liboracle_v4.so
    (AVX-512, the CPU baseline's build on hosts that have it)"""
    for extra in ([], ["OUT=liboracle_v4.so", "MARCH=x86-64-v4"]):
        proc = subprocess.run(["make", "-s", "-C", str(HERE)] + extra, capture_output=True, text=True)
        if proc.returncode != 0:
            raise RuntimeError(f"oracle build failed:\n{proc.stdout}\n{proc.stderr}")
    return LIB_PATH


def _load(path: Path) -> C.CDLL:
    lb = C.CDLL(str(path))
    lb.oracle_grid_hash.restype = C.c_uint32
    lb.oracle_dambreak_count.restype = C.c_uint32
    lb.oracle_peristalsis_mass_sum.restype = C.c_float
    lb.oracle_get_threads.restype = C.c_int
    return lb


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = _load(LIB_PATH)
    return _lib


def host_has_avx512() -> bool:
    try:
        flags = Path("/proc/cpuinfo").read_text()
    except OSError:
        return False
    return all(f" {f}" in flags for f in ("avx512f", "avx512bw", "avx512dq", "avx512vl"))


def use_default(path) -> None:
    """load the library at path (another ISA build of the same source) for
    every later call"""
    global _lib
    _lib = _load(Path(path))


def build_variant(mutation: int, outdir) -> Path:
    """The oracle built with the planted mistake ORACLE_MUTATE=mutation into
    outdir (a scratch directory): the pin tests check that they reject it."""
    out = Path(outdir).resolve() / f"liboracle_mut{int(mutation)}.so"
    proc = subprocess.run(["make", "-s", "-C", str(HERE), f"OUT={out}", f"DEFS=-DORACLE_MUTATE={int(mutation)}"],
                          capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"oracle variant build failed:\n{proc.stdout}\n{proc.stderr}")
    return out


@contextlib.contextmanager
def using(path):
    """Every function of this module calls the library at path inside the block."""
    global _lib
    saved = lib()
    _lib = _load(Path(path))
    try:
        yield
    finally:
        _lib = saved


def _ptr(a: np.ndarray | None):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _f(a, dtype):
    a = np.ascontiguousarray(a, dtype=dtype)
    return a


def set_threads(n: int) -> None:
    lib().oracle_set_threads(int(n))


def get_threads() -> int:
    return int(lib().oracle_get_threads())


def grid_pos(P, x: float, y: float, z: float) -> tuple[int, int, int]:
    out = (C.c_int32 * 3)()
    lib().oracle_grid_pos(C.byref(P), C.c_float(x), C.c_float(y), C.c_float(z), out)
    return tuple(out)


def grid_hash(P, g) -> int:
    arr = (C.c_int32 * 3)(*[int(v) for v in g])
    return int(lib().oracle_grid_hash(C.byref(P), arr))


def calc_hash(P, pos4: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    pos4 = _f(pos4, np.float32)
    n = pos4.shape[0]
    h = np.empty(n, np.uint32)
    idx = np.empty(n, np.uint32)
    lib().oracle_calc_hash(C.byref(P), _ptr(pos4), _ptr(h), _ptr(idx), C.c_uint32(n))
    return h, idx


def sort_by_key(keys: np.ndarray, vals: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    k = np.array(keys, dtype=np.uint32, copy=True)
    v = np.array(vals, dtype=np.uint32, copy=True)
    lib().oracle_sort_by_key(_ptr(k), _ptr(v), C.c_uint32(k.size))
    return k, v


def find_cell_start(sorted_hash: np.ndarray, num_cells: int,
                    cell_end_init: int = 0) -> tuple[np.ndarray, np.ndarray]:
    sh = _f(sorted_hash, np.uint32)
    cs = np.empty(num_cells, np.uint32)
    ce = np.full(num_cells, cell_end_init, np.uint32)
    lib().oracle_find_cell_start(_ptr(cs), _ptr(ce), _ptr(sh), C.c_uint32(sh.size), C.c_uint32(num_cells))
    return cs, ce


def _sfx(dtype) -> str:
    return "_f32" if np.dtype(dtype) == np.float32 else "_f64"


def gather(index, pos, vel, dtype=np.float32):
    # every converted array is bound to a name for the duration of the call:
    # a temporary's buffer would be freed before the C code reads it
    pos = _f(pos, dtype); vel = _f(vel, dtype); idx = _f(index, np.uint32)
    n = pos.shape[0]
    sp = np.empty_like(pos); sv = np.empty_like(vel)
    getattr(lib(), "oracle_gather" + _sfx(dtype))(_ptr(sp), _ptr(sv), _ptr(idx), _ptr(pos), _ptr(vel),
                                                 C.c_uint32(n))
    return sp, sv


def density(P, sorted_pos, cell_start, cell_end, measures_in=None, dtype=np.float32):
    sp = _f(sorted_pos, dtype)
    cs, ce = _f(cell_start, np.uint32), _f(cell_end, np.uint32)
    n = sp.shape[0]
    meas = np.zeros((n, 4), dtype) if measures_in is None else np.array(measures_in, dtype=dtype, copy=True)
    getattr(lib(), "oracle_density" + _sfx(dtype))(C.byref(P), _ptr(meas), _ptr(sp), _ptr(cs), _ptr(ce),
                                                  C.c_uint32(n))
    return meas


def force(P, measures, sorted_pos, sorted_vel, index, cell_start, cell_end, acc_in=None,
          dtype=np.float32):
    sp, m, sv = _f(sorted_pos, dtype), _f(measures, dtype), _f(sorted_vel, dtype)
    idx, cs, ce = _f(index, np.uint32), _f(cell_start, np.uint32), _f(cell_end, np.uint32)
    n = sp.shape[0]
    acc = np.zeros((n, 4), dtype) if acc_in is None else np.array(acc_in, dtype=dtype, copy=True)
    getattr(lib(), "oracle_force" + _sfx(dtype))(C.byref(P), _ptr(acc), _ptr(m), _ptr(sp), _ptr(sv), _ptr(idx),
                                                _ptr(cs), _ptr(ce), C.c_uint32(n))
    return acc


def pressure_force(P, measures, sorted_pos, index, cell_start, cell_end, dtype=np.float32):
    sp, m = _f(sorted_pos, dtype), _f(measures, dtype)
    idx, cs, ce = _f(index, np.uint32), _f(cell_start, np.uint32), _f(cell_end, np.uint32)
    n = sp.shape[0]
    out = np.zeros((n, 4), dtype)
    getattr(lib(), "oracle_pressure_force" + _sfx(dtype))(C.byref(P), _ptr(out), _ptr(m), _ptr(sp), _ptr(idx),
                                                         _ptr(cs), _ptr(ce), C.c_uint32(n))
    return out


def viscous_force(P, measures, sorted_pos, sorted_vel, index, cell_start, cell_end, dtype=np.float32):
    sp, m, sv = _f(sorted_pos, dtype), _f(measures, dtype), _f(sorted_vel, dtype)
    idx, cs, ce = _f(index, np.uint32), _f(cell_start, np.uint32), _f(cell_end, np.uint32)
    n = sp.shape[0]
    out = np.zeros((n, 4), dtype)
    getattr(lib(), "oracle_viscous_force" + _sfx(dtype))(C.byref(P), _ptr(out), _ptr(m), _ptr(sp), _ptr(sv),
                                                        _ptr(idx), _ptr(cs), _ptr(ce), C.c_uint32(n))
    return out


def integrate(P, pos, vel, vel_lf, acc, pres_f=None, wave_time=0.0, dtype=np.float32):
    pos = np.array(pos, dtype=dtype, copy=True)
    vel = np.array(vel, dtype=dtype, copy=True)
    vlf = np.array(vel_lf, dtype=dtype, copy=True)
    acc = _f(acc, dtype)
    pf = None if pres_f is None else _f(pres_f, dtype)
    R = C.c_float if np.dtype(dtype) == np.float32 else C.c_double
    getattr(lib(), "oracle_integrate" + _sfx(dtype))(C.byref(P), _ptr(pos), _ptr(vel), _ptr(vlf), _ptr(acc),
                                                    _ptr(pf), R(wave_time), C.c_uint32(pos.shape[0]))
    return pos, vel, vlf


def configure_boundary(P, pos, wave_height, dtype=np.float32):
    pos = np.array(pos, dtype=dtype, copy=True)
    R = C.c_float if np.dtype(dtype) == np.float32 else C.c_double
    getattr(lib(), "oracle_configure_boundary" + _sfx(dtype))(C.byref(P), _ptr(pos), R(wave_height),
                                                             C.c_uint32(pos.shape[0]))
    return pos


def shift_right_boundary(P, pos, dtype=np.float32):
    pos = np.array(pos, dtype=dtype, copy=True)
    getattr(lib(), "oracle_shift_right_boundary" + _sfx(dtype))(C.byref(P), _ptr(pos), C.c_uint32(pos.shape[0]))
    return pos


def remove_right_boundary(P, pos, dtype=np.float32):
    pos = np.array(pos, dtype=dtype, copy=True)
    getattr(lib(), "oracle_remove_right_boundary" + _sfx(dtype))(C.byref(P), _ptr(pos), C.c_uint32(pos.shape[0]))
    return pos


class _State32(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("pos", "vel", "velLF", "acc", "presF", "measures",
                                          "sortedPos", "sortedVel", "hash", "index", "cellStart",
                                          "cellEnd")] + \
               [("n", C.c_uint32), ("numCells", C.c_uint32),
                ("elapsed", C.c_float), ("time_shift", C.c_float), ("time_relax", C.c_float),
                ("wave_height", C.c_float)]


class _State64(C.Structure):
    _fields_ = _State32._fields_[:14] + [("elapsed", C.c_double), ("time_shift", C.c_double),
                                         ("time_relax", C.c_double), ("wave_height", C.c_double)]


class Simulation:
    """update() of the oracle over owned numpy arrays (one System object)."""

    def __init__(self, P, pos, vel=None, vel_lf=None, measures=None, dtype=np.float32):
        self.P = P.copy() if hasattr(P, "copy") else P
        self.dtype = np.dtype(dtype)
        n = np.asarray(pos).shape[0]
        self.n = n
        self.num_cells = int(P.grid_size[0]) * int(P.grid_size[1]) * int(P.grid_size[2])
        z = lambda: np.zeros((n, 4), dtype)  # noqa: E731
        self.pos = np.array(pos, dtype=dtype, copy=True)
        self.vel = z() if vel is None else np.array(vel, dtype=dtype, copy=True)
        self.velLF = z() if vel_lf is None else np.array(vel_lf, dtype=dtype, copy=True)
        self.acc = z()
        self.presF = z()
        if measures is None:
            self.measures = z()
            self.measures[:, 0] = P.rest_density
        else:
            self.measures = np.array(measures, dtype=dtype, copy=True)
        self.sortedPos = z()
        self.sortedVel = z()
        self.hash = np.zeros(n, np.uint32)
        self.index = np.zeros(n, np.uint32)
        self.cellStart = np.zeros(self.num_cells, np.uint32)
        self.cellEnd = np.zeros(self.num_cells, np.uint32)
        St = _State32 if self.dtype == np.float32 else _State64
        self.state = St()
        for name in ("pos", "vel", "velLF", "acc", "presF", "measures", "sortedPos", "sortedVel", "hash",
                     "index", "cellStart", "cellEnd"):
            setattr(self.state, name, getattr(self, name).ctypes.data)
        self.state.n = n
        self.state.numCells = self.num_cells
        self.state.elapsed = 0.0
        self.state.time_shift = 0.0
        self.state.time_relax = 1000 * float(np.float32(P.delta_time))
        self.state.wave_height = 0.0

    @property
    def elapsed(self) -> float:
        return float(self.state.elapsed)

    def update(self, steps: int = 1) -> None:
        fn = lib().oracle_update_f32 if self.dtype == np.float32 else lib().oracle_update_f64
        for _ in range(steps):
            fn(C.byref(self.P), C.byref(self.state))


def dambreak_init(P, n: int) -> np.ndarray:
    out = np.zeros((n, 4), np.float32)
    lib().oracle_dambreak_init(C.byref(P), _ptr(out), C.c_uint32(n))
    return out


def poiseuille_init(P, n: int) -> np.ndarray:
    out = np.zeros((n, 4), np.float32)
    lib().oracle_poiseuille_init(C.byref(P), _ptr(out), C.c_uint32(n))
    return out


def peristalsis_init(P, n: int) -> np.ndarray:
    out = np.zeros((n, 4), np.float32)
    lib().oracle_peristalsis_init(C.byref(P), _ptr(out), C.c_uint32(n))
    return out


def peristalsis_mass_sum(P, pos4) -> float:
    pos4 = _f(pos4, np.float32)
    return float(lib().oracle_peristalsis_mass_sum(C.byref(P), _ptr(pos4), C.c_uint32(pos4.shape[0])))


def dambreak_count(fluid, bo, grid) -> int:
    return int(lib().oracle_dambreak_count(*[C.c_uint32(int(v)) for v in fluid], C.c_int32(bo),
                                           *[C.c_uint32(int(v)) for v in grid]))


def run_pipeline(P, pos, vel_lf, measures=None, dtype=np.float32):
    """hash -> stable sort -> cellStart/End -> gather -> density, returning the
    intermediates (all in the reference's sorted order)."""
    pos32 = np.asarray(pos, np.float32)
    h, idx = calc_hash(P, pos32)
    sh, perm = sort_by_key(h, idx)
    nc = int(P.grid_size[0]) * int(P.grid_size[1]) * int(P.grid_size[2])
    cs, ce = find_cell_start(sh, nc)
    sp, sv = gather(perm, np.asarray(pos, dtype), np.asarray(vel_lf, dtype), dtype=dtype)
    meas = density(P, sp, cs, ce, measures_in=measures, dtype=dtype)
    return dict(hash=h, sorted_hash=sh, index=perm, cell_start=cs, cell_end=ce, sorted_pos=sp,
                sorted_vel=sv, measures=meas)


# ---------------------------------------------------------------- adjoint (Poiseuille)
def density_vjp(P, gpos_sorted, grho, sorted_pos, cell_start, cell_end, dtype=np.float64):
    """gpos_sorted += d(rho . grho)/d(sorted positions) (walls skipped)"""
    g = np.array(gpos_sorted, dtype=dtype, copy=True)
    gr, sp = _f(grho, dtype), _f(sorted_pos, dtype)
    cs, ce = _f(cell_start, np.uint32), _f(cell_end, np.uint32)
    getattr(lib(), "oracle_density_vjp" + _sfx(dtype))(C.byref(P), _ptr(g), _ptr(gr), _ptr(sp), _ptr(cs),
                                                      _ptr(ce), C.c_uint32(sp.shape[0]))
    return g


def force_vjp(P, gacc_sorted, measures, sorted_pos, sorted_vel, cell_start, cell_end, dtype=np.float64):
    """adjoints of sorted positions, leap-frog velocities, densities and the
    per-target share of d/d mu, from the sorted acceleration adjoint"""
    ga, m = _f(gacc_sorted, dtype), _f(measures, dtype)
    sp, sv = _f(sorted_pos, dtype), _f(sorted_vel, dtype)
    cs, ce = _f(cell_start, np.uint32), _f(cell_end, np.uint32)
    n = sp.shape[0]
    gpos = np.zeros((n, 4), dtype); gvel = np.zeros((n, 4), dtype)
    grho = np.zeros(n, dtype); gmu = np.zeros(n, dtype)
    getattr(lib(), "oracle_force_vjp" + _sfx(dtype))(C.byref(P), _ptr(gpos), _ptr(gvel), _ptr(grho), _ptr(gmu),
                                                    _ptr(ga), _ptr(m), _ptr(sp), _ptr(sv), _ptr(cs), _ptr(ce),
                                                    C.c_uint32(n))
    return gpos, gvel, grho, gmu


def integrate_vjp(P, pos, gpos, gvel, gvel_lf, ggrav, dtype=np.float64):
    """in place on (gpos, gvel, gvel_lf, ggrav); returns gacc"""
    p = _f(pos, dtype)
    n = p.shape[0]
    gacc = np.zeros((n, 4), dtype)
    for a in (gpos, gvel, gvel_lf, ggrav):
        assert a.dtype == np.dtype(dtype) and a.flags.c_contiguous
    getattr(lib(), "oracle_integrate_vjp" + _sfx(dtype))(C.byref(P), _ptr(p), _ptr(gpos), _ptr(gvel),
                                                        _ptr(gvel_lf), _ptr(gacc), _ptr(ggrav), C.c_uint32(n))
    return gacc


def profile_loss(P, pos, vel, target, bins: int = 64):
    """L = sum over non-empty y-bins of (mean x-velocity of the fluid in the
    bin - target)^2; bins split [wall_bottom, wall_top] evenly.  Returns
    (L, profile (nan where empty), dL/dvel (n,4))."""
    pos = np.asarray(pos, np.float64); vel = np.asarray(vel, np.float64)
    fl = pos[:, 3] == 0.0
    lo, hi = float(P.wall_bottom), float(P.wall_top)
    b = np.clip(np.floor((pos[fl, 1] - lo) / ((hi - lo) / bins)).astype(np.int64), 0, bins - 1)
    cnt = np.bincount(b, minlength=bins).astype(np.float64)
    s = np.bincount(b, weights=vel[fl, 0], minlength=bins)
    nz = cnt > 0
    prof = np.full(bins, np.nan)
    prof[nz] = s[nz] / cnt[nz]
    diff = np.where(nz, prof - np.asarray(target, np.float64), 0.0)
    L = float(np.sum(diff[nz] ** 2))
    g = np.zeros_like(vel)
    g[np.flatnonzero(fl), 0] = 2.0 * diff[b] / cnt[b]
    return L, prof, g


def rollout(P, pos, vel, vel_lf, steps: int, dtype=np.float64, checkpoint_every: int = 0):
    sim = Simulation(P, pos, vel, vel_lf, dtype=dtype)
    cps = {}
    for t in range(steps):
        if checkpoint_every and t % checkpoint_every == 0:
            cps[t] = (sim.pos.copy(), sim.vel.copy(), sim.velLF.copy())
        sim.update()
    return sim, cps


def poiseuille_adjoint(P, pos, vel, vel_lf, steps: int, target, bins: int = 64, checkpoint_every: int = 20,
                       dtype=np.float64) -> dict:
    """Reverse-mode gradient of profile_loss after `steps` updates with
    respect to mu (P.mu) and the body force P.gravity, by the hand adjoint:
    checkpoints every `checkpoint_every` steps, each segment recomputed
    forward and then walked backwards step by step."""
    sim, cps = rollout(P, pos, vel, vel_lf, steps, dtype, checkpoint_every)
    L, prof, gvel = profile_loss(P, sim.pos, sim.vel, target, bins)
    n = sim.n
    gvel = gvel.astype(dtype)
    gpos = np.zeros((n, 4), dtype); glf = np.zeros((n, 4), dtype); ggrav = np.zeros((n, 4), dtype)
    gmu = 0.0
    for c in sorted(cps, reverse=True):
        t1 = min(c + checkpoint_every, steps)
        s = Simulation(P, *cps[c], dtype=dtype)
        seg = []
        for t in range(c, t1):
            seg.append((s.pos.copy(), s.vel.copy(), s.velLF.copy()))
            if t < t1 - 1:
                s.update()
        for spos, svel, slf in reversed(seg):
            ref = run_pipeline(P, spos, slf, dtype=dtype)
            gacc = integrate_vjp(P, spos, gpos, gvel, glf, ggrav, dtype)
            idx = ref["index"].astype(np.int64)
            gp_s, gl_s, grho, gmu_s = force_vjp(P, gacc[idx], ref["measures"], ref["sorted_pos"],
                                                ref["sorted_vel"], ref["cell_start"], ref["cell_end"], dtype)
            gp_s = density_vjp(P, gp_s, grho, ref["sorted_pos"], ref["cell_start"], ref["cell_end"], dtype)
            gpos[idx] += gp_s
            glf[idx] = gl_s
            gmu += float(gmu_s.sum())
    return dict(loss=L, profile=prof, dmu=gmu, dgrav=ggrav[:, :3].sum(axis=0), gpos=gpos, gvel=gvel, gvel_lf=glf)


def density_subset(P, sorted_pos, cell_start, cell_end, sel, measures_in=None, dtype=np.float32):
    """density of the sorted particles sel only (rows of the others unchanged)"""
    sp = _f(sorted_pos, dtype)
    cs, ce, sl = _f(cell_start, np.uint32), _f(cell_end, np.uint32), _f(sel, np.uint32)
    n = sp.shape[0]
    meas = np.zeros((n, 4), dtype) if measures_in is None else np.array(measures_in, dtype=dtype, copy=True)
    getattr(lib(), "oracle_density_subset" + _sfx(dtype))(C.byref(P), _ptr(meas), _ptr(sp), _ptr(cs), _ptr(ce),
                                                         _ptr(sl), C.c_uint32(sl.size))
    return meas


def force_subset(P, measures, sorted_pos, sorted_vel, index, cell_start, cell_end, sel, acc_in=None,
                 dtype=np.float32):
    """force of the sorted particles sel only, scattered to acc[index[sel]]"""
    sp, m, sv = _f(sorted_pos, dtype), _f(measures, dtype), _f(sorted_vel, dtype)
    idx, cs, ce = _f(index, np.uint32), _f(cell_start, np.uint32), _f(cell_end, np.uint32)
    sl = _f(sel, np.uint32)
    n = sp.shape[0]
    acc = np.zeros((n, 4), dtype) if acc_in is None else np.array(acc_in, dtype=dtype, copy=True)
    getattr(lib(), "oracle_force_subset" + _sfx(dtype))(C.byref(P), _ptr(acc), _ptr(m), _ptr(sp), _ptr(sv),
                                                       _ptr(idx), _ptr(cs), _ptr(ce), _ptr(sl), C.c_uint32(sl.size))
    return acc
