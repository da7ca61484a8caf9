"""Batches of independent systems stepped together (cmag_batch_* of the C ABI).

    b = Batch([(sc.params, sc.pos, sc.vel) for sc in scenarios.channel_batch()])
    b.autotune()              # launch shape: block, neighbour-list depth, persistence
    b.run(1000)               # HIP graphs of tuning.steps_per_graph steps
    pos = b.getArray(k, ARRAY_POSITION)

Each instance steps exactly as a System built from the same parameter block and
state (cmag_amd.systems.ParamSystem) does, bit for bit.  The reference runs one
system per process (its parameters live in a process-global __constant__,
fluid_kernel.cu:17); a batch has no reference counterpart.
"""
from __future__ import annotations

import ctypes as C
import time
from typing import Iterable, Sequence

import numpy as np

from . import _lib as L
from ._lib import BatchTuning, CmagParams, call, lib


class Batch:
    def __init__(self, instances: Sequence[tuple]):
        """instances: (params, pos (n,4) float32, vel (n,4) float32 or None);
        velLF starts equal to vel"""
        m = len(instances)
        if m == 0:
            raise ValueError("a batch needs at least one instance")
        arr = (CmagParams * m)()
        counts = (C.c_uint32 * m)()
        for k, inst in enumerate(instances):
            arr[k] = inst[0]
            counts[k] = int(np.asarray(inst[1]).shape[0])
        h = C.c_void_p()
        call("cmag_batch_create", arr, counts, m, C.byref(h))
        self._h = h
        self.m = m
        self.counts = [int(c) for c in counts]
        for k, inst in enumerate(instances):
            pos = inst[1]
            vel = inst[2] if len(inst) > 2 and inst[2] is not None else np.zeros_like(pos)
            self.setArray(k, L.ARRAY_POSITION, pos)
            self.setArray(k, L.ARRAY_VELOCITY, vel)
            self.setArray(k, L.ARRAY_VELOCITYLEAPFROG, vel)

    # -- lifetime
    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().cmag_batch_destroy(self._h)
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

    # -- stepping
    def update(self) -> None:
        call("cmag_batch_update", self._h)

    def run(self, steps: int, use_graph: bool = True) -> None:
        call("cmag_batch_run", self._h, steps, 1 if use_graph else 0)

    def synchronize(self) -> None:
        call("cmag_batch_synchronize", self._h)

    def reset_clocks(self) -> None:
        call("cmag_batch_reset_clocks", self._h)

    # -- state
    def setArray(self, inst: int, array: int, data) -> None:
        a = np.ascontiguousarray(data, dtype=np.float32).reshape(-1, 4)
        call("cmag_batch_set_array", self._h, inst, array, a.ctypes.data, a.shape[0])

    def getArray(self, inst: int, array: int) -> np.ndarray:
        out = np.empty((self.counts[inst], 4), np.float32)
        call("cmag_batch_get_array", self._h, inst, array, out.ctypes.data, out.shape[0])
        return out

    def instance_range(self, inst: int) -> tuple[int, int]:
        first, count = C.c_uint32(), C.c_uint32()
        call("cmag_batch_instance_range", self._h, inst, C.byref(first), C.byref(count))
        return int(first.value), int(count.value)

    def params(self, inst: int) -> CmagParams:
        p = CmagParams()
        call("cmag_batch_get_params", self._h, inst, C.byref(p))
        return p

    def elapsed(self, inst: int) -> float:
        t = C.c_float()
        call("cmag_batch_elapsed_time", self._h, inst, C.byref(t))
        return float(t.value)

    def device_ptr(self, array: int) -> int:
        return lib().cmag_batch_device_ptr(self._h, array) or 0

    @property
    def num_particles(self) -> int:
        return int(lib().cmag_batch_num_particles(self._h))

    @property
    def num_cells(self) -> int:
        return int(lib().cmag_batch_num_cells(self._h))

    # -- launch shape
    @property
    def tuning(self) -> dict:
        t = BatchTuning()
        call("cmag_batch_get_tuning", self._h, C.byref(t))
        return t.as_dict()

    def set_tuning(self, **kw) -> None:
        cur = self.tuning
        cur.update(kw)
        t = BatchTuning(**cur)
        call("cmag_batch_set_tuning", self._h, C.byref(t))

    def autotune(self, steps_per_trial: int = 5) -> tuple[dict, float]:
        """times every launch shape on this batch's own state (restored after
        each trial) and keeps the fastest; returns (tuning, ms per step)"""
        t = BatchTuning()
        ms = C.c_float()
        call("cmag_batch_autotune", self._h, steps_per_trial, C.byref(t), C.byref(ms))
        return t.as_dict(), float(ms.value)

    # -- timing
    def enable_timing(self, on: bool = True) -> None:
        call("cmag_batch_enable_timing", self._h, 1 if on else 0)

    def stage_times(self) -> dict:
        ms = (C.c_double * 6)()
        cnt = (C.c_uint32 * 6)()
        call("cmag_batch_stage_times", self._h, ms, cnt, 6)
        return {name: (ms[k], cnt[k]) for k, name in enumerate(L.STAGES)}


def autotune_cells(make_instances, cell_factors: Iterable[float] = (1.0, 1.3, 2.0), steps: int = 5,
                   steps_per_trial: int = 5) -> dict:
    """The cell-size axis of the autotuner: for each cell factor (cell >=
    factor * dx) builds the batch from make_instances(cell_factor), lets
    cmag_batch_autotune pick its launch shape, times `steps` graph steps, and
    returns the fastest {"cell_factor", "tuning", "ms_per_step", "trials"}.
    Different cell sizes change each instance's stencil walk (its sum order),
    not its physics."""
    trials = []
    for cf in cell_factors:
        b = Batch(make_instances(cf))
        try:
            tune, _ = b.autotune(steps_per_trial)
            b.run(1)
            b.synchronize()
            t0 = time.perf_counter()
            b.run(steps)
            b.synchronize()
            ms = 1e3 * (time.perf_counter() - t0) / steps
            trials.append({"cell_factor": cf, "tuning": tune, "ms_per_step": ms})
        finally:
            b.close()
    best = min(trials, key=lambda d: d["ms_per_step"])
    return dict(best, trials=trials)


__all__ = ["Batch", "autotune_cells"]
