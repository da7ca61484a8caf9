"""Python mirror of the reference System classes over the C ABI.

    DamBreakSystem(fluidParticlesSize, boundaryOffset, gridSize, particleRadius, bUseOpenGL)
        DamBreak.Core/fluidSystem.h:6-113          (FluidSystem is an alias)
    PoiseuilleFlowSystem(same arguments)
        Poiseuille.Core/poiseuilleFlowSystem.h:6-112
    PeristalsisSystem(deltaTime, fluid_size, amplitude, wave_speed, soundspeed,
                      gravity, boundaryOffset, gridSize, radius, bUseOpenGL)
        Peristalsis.Core/PeristalsisSystem.h:6-132

Method names, enums and argument meaning follow the reference headers;
getArray/update(dt) are the additions the C++ classes also carry.  Every call
goes through libcmag_hip.so; errors raise CmagError instead of exiting.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import CmagParams, call, lib


def _u3(v) -> tuple[int, int, int]:
    v = tuple(int(x) for x in v)
    if len(v) != 3:
        raise ValueError("expected three components")
    return v


class SystemBase:
    """Handle over a cmag_system; owns its device buffers."""

    # reference enum value -> CMAG_ARRAY_*; subclasses override
    ARRAYS: dict[int, int] = {}

    def __init__(self, handle: C.c_void_p):
        self._h = handle

    # -- lifetime
    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().cmag_system_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- reference API
    def update(self, dt: float | None = None) -> None:
        if dt is None:
            call("cmag_system_update", self._h)
        else:
            call("cmag_system_update_dt", self._h, C.c_float(dt))

    def reset(self) -> None:
        call("cmag_system_reset", self._h)

    def run(self, steps: int, use_graph: bool = True) -> None:
        call("cmag_system_run", self._h, steps, 1 if use_graph else 0)

    def _array_id(self, array: int) -> int:
        return self.ARRAYS.get(array, array) if self.ARRAYS else array

    def setArray(self, array: int, data, start: int = 0, count: int | None = None) -> None:
        a = np.ascontiguousarray(np.asarray(data, dtype=np.float32).reshape(-1, 4))
        if count is None:
            count = a.shape[0]
        if a.shape[0] < count:
            raise ValueError("data holds fewer particles than count")
        call("cmag_system_set_array", self._h, self._array_id(array), a.ctypes.data, start, count)

    def getArray(self, array: int, start: int = 0, count: int | None = None) -> np.ndarray:
        if count is None:
            count = self.getNumParticles() - start
        out = np.empty((count, 4), dtype=np.float32)
        call("cmag_system_get_array", self._h, self._array_id(array), out.ctypes.data, start, count)
        return out

    def getUintArray(self, array: int) -> np.ndarray:
        n = self.getNumParticles() if array in (L.ARRAY_HASH, L.ARRAY_INDEX) else self.num_cells
        out = np.empty(n, dtype=np.uint32)
        call("cmag_system_get_uint_array", self._h, array, out.ctypes.data, 0, n)
        return out

    def device_ptr(self, array: int) -> int:
        return lib().cmag_system_device_ptr(self._h, array) or 0

    def getNumParticles(self) -> int:
        return int(lib().cmag_system_num_particles(self._h))

    @property
    def num_cells(self) -> int:
        return int(lib().cmag_system_num_cells(self._h))

    def getElapsedTime(self) -> float:
        return float(lib().cmag_system_elapsed_time(self._h))

    @property
    def params(self) -> CmagParams:
        p = CmagParams()
        call("cmag_system_get_params", self._h, C.byref(p))
        return p

    def set_params(self, p: CmagParams) -> None:
        call("cmag_system_set_params", self._h, C.byref(p))

    def enable_timing(self, on: bool = True) -> None:
        call("cmag_system_enable_timing", self._h, 1 if on else 0)

    def stage_times(self) -> dict:
        """{stage: (total ms, launches)} since the last call; clears the totals"""
        ms = (C.c_double * 6)()
        cnt = (C.c_uint32 * 6)()
        call("cmag_system_stage_times", self._h, ms, cnt, 6)
        return {name: (ms[k], cnt[k]) for k, name in enumerate(L.STAGES)}

    def set_subcell_walk(self, on: bool) -> None:
        """DamBreak neighbour passes over the internal sub-cell order (on by
        default where it applies; results agree with the plain walk to rounding)"""
        call("cmag_system_set_subcell_walk", self._h, 1 if on else 0)

    @property
    def subcell_walk(self) -> bool:
        return bool(lib().cmag_system_subcell_walk(self._h))

    def set_inrange_masks(self, on: bool) -> None:
        """the force pass takes wall-free rows' in-range neighbours from the
        density pass's masks (on by default; forces identical either way)"""
        call("cmag_system_set_inrange_masks", self._h, 1 if on else 0)

    @property
    def inrange_masks(self) -> bool:
        return bool(lib().cmag_system_inrange_masks(self._h))

    def synchronize(self) -> None:
        call("cmag_system_synchronize", self._h)


    # device-side state access (stream-ordered on the system's stream)
    def copy_from_device(self, array: int, device_ptr: int, count: int | None = None) -> None:
        """rows [0, count) of `array` <- device memory at device_ptr"""
        n = self.getNumParticles() if count is None else count
        call("cmag_system_copy_array_device", self._h, array, device_ptr, 0, n, 1)

    def copy_to_device(self, array: int, device_ptr: int, count: int | None = None) -> None:
        """device memory at device_ptr <- rows [0, count) of `array`"""
        n = self.getNumParticles() if count is None else count
        call("cmag_system_copy_array_device", self._h, array, device_ptr, 0, n, 0)

    def vjp_step(self, gpos: int, gvel: int, gvel_lf: int, gmu_acc: int = 0, ggrav_acc: int = 0) -> None:
        """adjoint of one update() at the current state (device pointers; see
        cmag_system_vjp_step)"""
        call("cmag_system_vjp_step", self._h, gpos, gvel, gvel_lf, gmu_acc or None, ggrav_acc or None)

    def set_stream(self, stream: int) -> None:
        call("cmag_system_set_stream", self._h, stream)

    # geometry getters shared by the three headers
    def getHalfWorldXSize(self) -> float:
        p = self.params
        return np.float32(p.grid_size[0]) * np.float32(p.particle_radius)

    def getHalfWorldYSize(self) -> float:
        p = self.params
        return np.float32(p.grid_size[1]) * np.float32(p.particle_radius)

    def getHalfWorldZSize(self) -> float:
        p = self.params
        return np.float32(p.grid_size[2]) * np.float32(p.particle_radius)

    def getGridSize(self) -> tuple[int, int, int]:
        return tuple(self.params.grid_size)

    def getWorldOrigin(self) -> tuple[float, float, float]:
        return tuple(self.params.world_origin)

    def getCellSize(self) -> tuple[float, float, float]:
        return tuple(self.params.cell_size)

    # raw device pointers (borrowed), fluidSystem.h:40-47
    def getCudaPosVBO(self) -> int:
        return self.device_ptr(L.ARRAY_POSITION)

    def getCudaVelVBO(self) -> int:
        return self.device_ptr(L.ARRAY_VELOCITY)

    def getCudaHash(self) -> int:
        return self.device_ptr(L.ARRAY_HASH)

    def getCudaIndex(self) -> int:
        return self.device_ptr(L.ARRAY_INDEX)

    def getCudaSortedPosition(self) -> int:
        return self.device_ptr(L.ARRAY_SORTEDPOS)


class DamBreakSystem(SystemBase):
    POSITION, VELOCITY, MEASURES, ACCELERATION, VELOCITYLEAPFROG = range(5)
    ARRAYS = {0: L.ARRAY_POSITION, 1: L.ARRAY_VELOCITY, 2: L.ARRAY_MEASURES,
              3: L.ARRAY_ACCELERATION, 4: L.ARRAY_VELOCITYLEAPFROG}

    def __init__(self, fluidParticlesSize, boundaryOffset: int, gridSize, particleRadius: float,
                 bUseOpenGL: bool = False):
        h = C.c_void_p()
        fx, fy, fz = _u3(fluidParticlesSize)
        gx, gy, gz = _u3(gridSize)
        call("cmag_dambreak_create", fx, fy, fz, boundaryOffset, gx, gy, gz, C.c_float(particleRadius),
             C.byref(h))
        super().__init__(h)

    def getCudaMeasures(self) -> int:
        return self.device_ptr(L.ARRAY_MEASURES)

    def getCudaAcceleration(self) -> int:
        return self.device_ptr(L.ARRAY_ACCELERATION)

    def changeRightBoundary(self) -> None:
        call("cmag_system_change_right_boundary", self._h)

    def removeRightBoundary(self) -> None:
        call("cmag_system_remove_right_boundary", self._h)

    def getParticleRadius(self) -> float:
        return self.params.particle_radius

    def getGravity(self) -> tuple[float, float, float]:
        return tuple(self.params.gravity)


FluidSystem = DamBreakSystem


class PoiseuilleFlowSystem(SystemBase):
    POSITION, VELOCITY, MEASURES, ACCELERATION, VELOCITYLEAPFROG = range(5)
    ARRAYS = DamBreakSystem.ARRAYS

    def __init__(self, fluidParticlesSize, boundaryOffset: int, gridSize, particleRadius: float,
                 bUseOpenGL: bool = False):
        h = C.c_void_p()
        fx, fy, fz = _u3(fluidParticlesSize)
        gx, gy, gz = _u3(gridSize)
        call("cmag_poiseuille_create", fx, fy, fz, boundaryOffset, gx, gy, gz, C.c_float(particleRadius),
             C.byref(h))
        super().__init__(h)

    def getCudaMeasures(self) -> int:
        return self.device_ptr(L.ARRAY_MEASURES)

    def getCudaAcceleration(self) -> int:
        return self.device_ptr(L.ARRAY_ACCELERATION)

    def getLeapFrogVelocity(self) -> int:
        return self.device_ptr(L.ARRAY_VELOCITYLEAPFROG)

    def changeGravity(self) -> None:
        call("cmag_system_change_gravity", self._h)

    def getParticleRadius(self) -> float:
        return self.params.particle_radius


class PeristalsisSystem(SystemBase):
    (POSITION, PREDICTEDPOSITION, VELOCITY, MEASURES, VISCOUSFORCE, PRESSUREFORCE,
     VELOCITYLEAPFROG) = range(7)
    ARRAYS = {0: L.ARRAY_POSITION, 1: L.ARRAY_PREDICTEDPOSITION, 2: L.ARRAY_VELOCITY,
              3: L.ARRAY_MEASURES, 4: L.ARRAY_ACCELERATION, 5: L.ARRAY_PRESSUREFORCE,
              6: L.ARRAY_VELOCITYLEAPFROG}

    def __init__(self, deltaTime: float, fluid_size, amplitude: float, wave_speed: float,
                 soundspeed: float, gravity, boundaryOffset: int, gridSize, radius: float,
                 bUseOpenGL: bool = False):
        h = C.c_void_p()
        fx, fy, fz = _u3(fluid_size)
        gx, gy, gz = _u3(gridSize)
        g = [C.c_float(float(v)) for v in gravity]
        call("cmag_peristalsis_create", C.c_float(deltaTime), fx, fy, fz, C.c_float(amplitude),
             C.c_float(wave_speed), C.c_float(soundspeed), *g, boundaryOffset, gx, gy, gz,
             C.c_float(radius), C.byref(h))
        super().__init__(h)

    def Update(self) -> None:
        self.update()

    def Reset(self) -> None:
        self.reset()

    def GetElapsedTime(self) -> float:
        return self.getElapsedTime()

    def getMeasures(self) -> int:
        return self.device_ptr(L.ARRAY_MEASURES)

    def viscous_force(self) -> int:
        return self.device_ptr(L.ARRAY_ACCELERATION)

    def pressure_force(self) -> int:
        return self.device_ptr(L.ARRAY_PRESSUREFORCE)

    def getSortedVelocity(self) -> int:
        return self.device_ptr(L.ARRAY_SORTEDVEL)

    def getPredictedPos(self) -> int:
        return self.device_ptr(L.ARRAY_PREDICTEDPOSITION)

    def getLeapFrogVelocity(self) -> int:
        return self.device_ptr(L.ARRAY_VELOCITYLEAPFROG)

    def getradius(self) -> float:
        return self.params.particle_radius


class ParamSystem(SystemBase):
    """A System on an arbitrary parameter block (configuration scenarios).

    Arrays use the CMAG_ARRAY_* ids directly; there is no reset()."""

    POSITION, VELOCITY, MEASURES, ACCELERATION, VELOCITYLEAPFROG = range(5)
    PRESSUREFORCE = L.ARRAY_PRESSUREFORCE

    def __init__(self, params: CmagParams, num_particles: int):
        h = C.c_void_p()
        call("cmag_system_create", C.byref(params), num_particles, C.byref(h))
        super().__init__(h)

    @classmethod
    def from_state(cls, params: CmagParams, pos4: np.ndarray, vel4: np.ndarray | None = None,
                   measures4: np.ndarray | None = None) -> "ParamSystem":
        n = int(pos4.shape[0])
        s = cls(params, n)
        s.setArray(L.ARRAY_POSITION, pos4)
        if vel4 is not None:
            s.setArray(L.ARRAY_VELOCITY, vel4)
            s.setArray(L.ARRAY_VELOCITYLEAPFROG, vel4)
        meas = np.zeros((n, 4), np.float32)
        meas[:, 0] = params.rest_density
        s.setArray(L.ARRAY_MEASURES, meas if measures4 is None else measures4)
        return s

    # the scenario controls of the reference classes, for parameter-block systems
    def changeRightBoundary(self) -> None:
        """DamBreak: gate and clamp move by the column width (fluidSystem.cpp:239-255)"""
        call("cmag_system_change_right_boundary", self._h)

    def removeRightBoundary(self) -> None:
        """DamBreak: gate parked, clamp lifted, clock restarted (fluidSystem.cpp:222-237)"""
        call("cmag_system_remove_right_boundary", self._h)

    def changeGravity(self) -> None:
        """flip gravity's y component (poiseuilleFlowSystem.cpp:207-210)"""
        call("cmag_system_change_gravity", self._h)
