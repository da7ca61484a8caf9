"""ctypes binding of libcmag_hip.so (the C ABI declared in include/cmag_hip.h).

The library is loaded from the package directory only; a missing library is a
hard error (there is no fallback path).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libcmag_hip.so"

CMAG_OK = 0
CMAG_E_INVALID, CMAG_E_HIP, CMAG_E_SIZE, CMAG_E_STATE = -1, -2, -3, -4


class CmagError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed ({code}): {msg}")
        self.code = code


class CmagParams(C.Structure):
    """Mirror of struct cmag_params (include/cmag_params.h); every field is 4 bytes."""

    _fields_ = [
        ("version", C.c_uint32),
        ("system", C.c_int32),
        ("eos", C.c_int32),
        ("hash_order", C.c_int32),
        ("gridpos_rule", C.c_int32),
        ("dim", C.c_int32),
        ("periodic_mask", C.c_int32),
        ("cellcount", C.c_int32),
        ("grid_size", C.c_uint32 * 3),
        ("grid_bits", C.c_uint32 * 3),
        ("fluid_size", C.c_uint32 * 3),
        ("world_origin", C.c_float * 3),
        ("cell_size", C.c_float * 3),
        ("world_size", C.c_float * 3),
        ("gravity", C.c_float * 3),
        ("particle_radius", C.c_float),
        ("smoothing_radius", C.c_float),
        ("particle_mass", C.c_float),
        ("rest_density", C.c_float),
        ("soundspeed", C.c_float),
        ("B", C.c_float),
        ("gamma", C.c_float),
        ("delta_time", C.c_float),
        ("boundary_damping", C.c_float),
        ("mu", C.c_float),
        ("alpha", C.c_float),
        ("kernel_coeff", C.c_float),
        ("grad_coeff", C.c_float),
        ("boundary_offset", C.c_int32),
        ("right_boundary", C.c_float),
        ("lj_D", C.c_float),
        ("lj_a", C.c_float),
        ("clamp_lo", C.c_float * 3),
        ("clamp_hi", C.c_float * 3),
        ("wall_bottom", C.c_float),
        ("wall_top", C.c_float),
        ("beta_max", C.c_float),
        ("amplitude", C.c_float),
        ("wave_speed", C.c_float),
        ("wave_k", C.c_float),
        ("is_boundary_configuration", C.c_int32),
        ("wall_velocity", C.c_int32),
        ("morton_mask", C.c_uint32 * 3),
        ("morton_xbits", C.c_uint32),
        ("reserved", C.c_int32 * 4),
    ]

    def copy(self) -> "CmagParams":
        out = CmagParams()
        C.memmove(C.byref(out), C.byref(self), C.sizeof(self))
        return out

    def as_dict(self) -> dict:
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            d[name] = list(v) if hasattr(v, "__len__") else v
        return d


# enum values (include/cmag_params.h, include/cmag_hip.h)
SYS_DAMBREAK, SYS_POISEUILLE, SYS_PERISTALSIS = 0, 1, 2
EOS_TAIT, EOS_LINEAR_C2, EOS_LINEAR_50C = 0, 1, 2
HASH_LINEAR, HASH_MORTON = 0, 1
GRIDPOS_DIV_CELL, GRIDPOS_HALF_DIV_RADIUS = 0, 1
ARRAY_POSITION = 0
ARRAY_VELOCITY = 1
ARRAY_MEASURES = 2
ARRAY_ACCELERATION = 3
ARRAY_VELOCITYLEAPFROG = 4
ARRAY_PRESSUREFORCE = 5
ARRAY_SORTEDPOS = 6
ARRAY_SORTEDVEL = 7
ARRAY_PREDICTEDPOSITION = 8
ARRAY_HASH = 16
ARRAY_INDEX = 17
ARRAY_CELLSTART = 18
ARRAY_CELLEND = 19
CELL_EMPTY = 0xFFFFFFFF
STAGES = ("hash", "sort", "reorder", "density", "force", "integrate")

class BatchTuning(C.Structure):
    """cmag_batch_tuning (include/cmag_hip.h)"""
    _fields_ = [("block", C.c_uint32), ("nb_cap", C.c_uint32), ("blocks_per_cu", C.c_uint32),
                ("steps_per_graph", C.c_uint32)]

    def as_dict(self) -> dict:
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


_P = C.POINTER(CmagParams)
_T = C.POINTER(BatchTuning)
_vp = C.c_void_p
_u32 = C.c_uint32
_i32 = C.c_int32
_f32 = C.c_float

# name -> (restype, argtypes); restype int means a status to check
SIGNATURES: dict[str, tuple] = {
    "cmag_last_error": (C.c_char_p, []),
    "cmag_version": (C.c_int, []),
    "cmag_params_size": (C.c_int, []),
    "cmag_params_dambreak": (C.c_int, [_P, _u32, _u32, _u32, _i32, _u32, _u32, _u32, _f32]),
    "cmag_params_poiseuille": (C.c_int, [_P, _u32, _u32, _u32, _i32, _u32, _u32, _u32, _f32]),
    "cmag_params_peristalsis": (C.c_int, [_P, _f32, _u32, _u32, _u32, _f32, _f32, _f32, _f32, _f32,
                                          _f32, _i32, _u32, _u32, _u32, _f32]),
    "cmag_params_finalize": (C.c_int, [_P]),
    "cmag_params_num_cells": (C.c_uint32, [_P]),
    "cmag_params_key_bits": (C.c_int, [_P]),
    "cmag_reference_initial_state": (C.c_int, [_P, _u32, _vp]),
    "cmag_peristalsis_lattice_sum": (C.c_float, [_P, _vp, _u32]),
    "cmag_calc_hash": (C.c_int, [_P, _vp, _vp, _vp, _u32, _vp]),
    "cmag_sort_scratch_bytes": (C.c_size_t, [_u32, C.c_int]),
    "cmag_sort_by_key": (C.c_int, [_vp, _vp, _u32, C.c_int, _vp, C.c_size_t, _vp]),
    "cmag_sort_noalloc_scratch_bytes": (C.c_size_t, []),
    "cmag_system_color_buffer": (C.c_int, [_vp, C.POINTER(_vp)]),
    "cmag_system_coloring": (C.c_int, [_vp]),
    "cmag_fill_colors": (C.c_int, [_P, _vp, _vp, _u32, _vp]),
    "cmag_render_view_default": (None, [_vp]),
    "cmag_renderer_create": (C.c_int, [_u32, _u32, C.POINTER(_vp)]),
    "cmag_renderer_destroy": (C.c_int, [_vp]),
    "cmag_renderer_display": (C.c_int, [_vp, _vp, _vp, _vp, _u32, _vp]),
    "cmag_renderer_image": (_vp, [_vp]),
    "cmag_renderer_read": (C.c_int, [_vp, _vp, _vp]),
    "cmag_sort_by_key_noalloc": (C.c_int, [_vp, _vp, _u32, C.c_int, _vp]),
    "cmag_reorder_find_cell_start": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _u32, _u32, _vp]),
    "cmag_density": (C.c_int, [_P, _vp, _vp, _vp, _vp, _u32, _vp]),
    "cmag_count_neighbours": (C.c_int, [_P, _vp, _vp, _vp, _vp, _u32, _vp]),
    "cmag_force": (C.c_int, [_P, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _u32, _vp]),
    "cmag_pressure_force": (C.c_int, [_P, _vp, _vp, _vp, _vp, _vp, _vp, _u32, _vp]),
    "cmag_viscous_force": (C.c_int, [_P, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _u32, _vp]),
    "cmag_integrate": (C.c_int, [_P, _vp, _vp, _vp, _vp, _vp, _f32, _u32, _vp]),
    "cmag_configure_boundary": (C.c_int, [_P, _vp, _f32, _u32, _vp]),
    "cmag_shift_right_boundary": (C.c_int, [_P, _vp, _u32, _vp]),
    "cmag_remove_right_boundary": (C.c_int, [_P, _vp, _u32, _vp]),
    "cmag_dambreak_create": (C.c_int, [_u32, _u32, _u32, _i32, _u32, _u32, _u32, _f32, C.POINTER(_vp)]),
    "cmag_poiseuille_create": (C.c_int, [_u32, _u32, _u32, _i32, _u32, _u32, _u32, _f32, C.POINTER(_vp)]),
    "cmag_peristalsis_create": (C.c_int, [_f32, _u32, _u32, _u32, _f32, _f32, _f32, _f32, _f32, _f32,
                                          _i32, _u32, _u32, _u32, _f32, C.POINTER(_vp)]),
    "cmag_system_create": (C.c_int, [_P, _u32, C.POINTER(_vp)]),
    "cmag_system_destroy": (C.c_int, [_vp]),
    "cmag_system_reset": (C.c_int, [_vp]),
    "cmag_system_update": (C.c_int, [_vp]),
    "cmag_system_update_dt": (C.c_int, [_vp, _f32]),
    "cmag_system_run": (C.c_int, [_vp, _u32, C.c_int]),
    "cmag_system_set_array": (C.c_int, [_vp, C.c_int, _vp, C.c_int, C.c_int]),
    "cmag_system_get_array": (C.c_int, [_vp, C.c_int, _vp, C.c_int, C.c_int]),
    "cmag_system_get_uint_array": (C.c_int, [_vp, C.c_int, _vp, _u32, _u32]),
    "cmag_system_device_ptr": (_vp, [_vp, C.c_int]),
    "cmag_system_num_particles": (C.c_uint32, [_vp]),
    "cmag_system_num_cells": (C.c_uint32, [_vp]),
    "cmag_system_elapsed_time": (C.c_float, [_vp]),
    "cmag_system_get_params": (C.c_int, [_vp, _P]),
    "cmag_system_set_params": (C.c_int, [_vp, _P]),
    "cmag_system_set_stream": (C.c_int, [_vp, _vp]),
    "cmag_system_set_subcell_walk": (C.c_int, [_vp, C.c_int]),
    "cmag_system_subcell_walk": (C.c_int, [_vp]),
    "cmag_system_set_inrange_masks": (C.c_int, [_vp, C.c_int]),
    "cmag_system_inrange_masks": (C.c_int, [_vp]),
    "cmag_system_synchronize": (C.c_int, [_vp]),
    "cmag_system_enable_timing": (C.c_int, [_vp, C.c_int]),
    "cmag_system_stage_times": (C.c_int, [_vp, C.POINTER(C.c_double), C.POINTER(C.c_uint32), C.c_int]),
    "cmag_system_change_right_boundary": (C.c_int, [_vp]),
    "cmag_system_remove_right_boundary": (C.c_int, [_vp]),
    "cmag_system_change_gravity": (C.c_int, [_vp]),
    "cmag_integrate_vjp": (C.c_int, [_P, _vp, _vp, _vp, _vp, _vp, _vp, _u32, _vp]),
    "cmag_force_vjp": (C.c_int, [_P, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _u32, _vp]),
    "cmag_density_vjp": (C.c_int, [_P, _vp, _vp, _vp, _vp, _vp, _u32, _vp]),
    "cmag_system_vjp_step": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "cmag_system_copy_array_device": (C.c_int, [_vp, C.c_int, _vp, C.c_int, C.c_int, C.c_int]),
    "cmag_batch_create": (C.c_int, [_P, C.POINTER(_u32), _u32, C.POINTER(_vp)]),
    "cmag_batch_destroy": (C.c_int, [_vp]),
    "cmag_batch_num_instances": (C.c_uint32, [_vp]),
    "cmag_batch_num_particles": (C.c_uint32, [_vp]),
    "cmag_batch_num_cells": (C.c_uint32, [_vp]),
    "cmag_batch_instance_range": (C.c_int, [_vp, _u32, C.POINTER(_u32), C.POINTER(_u32)]),
    "cmag_batch_get_params": (C.c_int, [_vp, _u32, _P]),
    "cmag_batch_set_array": (C.c_int, [_vp, _u32, C.c_int, _vp, _u32]),
    "cmag_batch_get_array": (C.c_int, [_vp, _u32, C.c_int, _vp, _u32]),
    "cmag_batch_device_ptr": (_vp, [_vp, C.c_int]),
    "cmag_batch_elapsed_time": (C.c_int, [_vp, _u32, C.POINTER(_f32)]),
    "cmag_batch_reset_clocks": (C.c_int, [_vp]),
    "cmag_batch_update": (C.c_int, [_vp]),
    "cmag_batch_run": (C.c_int, [_vp, _u32, C.c_int]),
    "cmag_batch_synchronize": (C.c_int, [_vp]),
    "cmag_batch_get_tuning": (C.c_int, [_vp, _T]),
    "cmag_batch_set_tuning": (C.c_int, [_vp, _T]),
    "cmag_batch_autotune": (C.c_int, [_vp, _u32, _T, C.POINTER(_f32)]),
    "cmag_batch_enable_timing": (C.c_int, [_vp, C.c_int]),
    "cmag_batch_stage_times": (C.c_int, [_vp, C.POINTER(C.c_double), C.POINTER(C.c_uint32), C.c_int]),
}

# entry points whose int return is a status
STATUS_FUNCS = {n for n, (r, _) in SIGNATURES.items()
                if r is C.c_int and n not in ("cmag_version", "cmag_params_size", "cmag_params_key_bits")}

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is not built; run `python -m cmag_amd.build` (HIP toolchain required)")
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        if handle.cmag_params_size() != C.sizeof(CmagParams):
            raise RuntimeError("cmag_params layout mismatch between libcmag_hip.so and the binding")
        _lib = handle
    return _lib


def last_error() -> str:
    msg = lib().cmag_last_error()
    return msg.decode() if msg else ""


def call(name: str, *args):
    """Calls an entry point; raises CmagError on a non-zero status."""
    rc = getattr(lib(), name)(*args)
    if name in STATUS_FUNCS and rc != CMAG_OK:
        raise CmagError(name, rc, last_error())
    return rc
