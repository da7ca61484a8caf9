"""Device-buffer helpers for the GPU tests: torch tensors as plain device memory,
stage launches through the C ABI, results back to numpy."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from cmag_amd import _lib as L


def to_dev(a: np.ndarray) -> torch.Tensor:
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        a = a.view(np.int32)
    return torch.from_numpy(a.copy()).to("cuda")


def empty_f4(n: int) -> torch.Tensor:
    return torch.zeros((max(n, 1), 4), dtype=torch.float32, device="cuda")


def empty_u32(n: int, fill: int | None = None) -> torch.Tensor:
    t = torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")
    if fill is not None:
        t.fill_(np.uint32(fill).view(np.int32).item())
    return t


def to_np(t: torch.Tensor, dtype=None) -> np.ndarray:
    torch.cuda.synchronize()
    a = t.detach().cpu().numpy()
    if t.dtype == torch.int32 and dtype is None:
        return a.view(np.uint32)
    return a


def P(p) -> C.c_void_p:
    return C.byref(p)


def ptr(t: torch.Tensor) -> int:
    return t.data_ptr()


class Stages:
    """C-ABI launchers over torch buffers (default stream, synchronised).

    Every device buffer is bound to a local name until the launch has been
    synchronised: a temporary's data_ptr would go back to the caching
    allocator at once and the next allocation could reuse its memory."""

    @staticmethod
    def calc_hash(p, pos4: np.ndarray):
        n = pos4.shape[0]
        d = to_dev(pos4.astype(np.float32))
        h, i = empty_u32(n), empty_u32(n)
        L.call("cmag_calc_hash", P(p), ptr(d), ptr(h), ptr(i), n, None)
        return to_np(h)[:n], to_np(i)[:n]

    @staticmethod
    def sort(keys: np.ndarray, vals: np.ndarray, key_bits: int):
        n = keys.size
        k, v = to_dev(keys.astype(np.uint32)), to_dev(vals.astype(np.uint32))
        nbytes = L.lib().cmag_sort_scratch_bytes(n, key_bits)
        scratch = torch.empty(max(nbytes, 1), dtype=torch.uint8, device="cuda")
        L.call("cmag_sort_by_key", ptr(k), ptr(v), n, key_bits, ptr(scratch), nbytes, None)
        out = to_np(k)[:n], to_np(v)[:n]
        del scratch
        return out

    @staticmethod
    def reorder(sorted_hash, index, pos, vel, num_cells, cell_end_fill=0xDEADBEEF):
        n = sorted_hash.size
        cs, ce = empty_u32(num_cells), empty_u32(num_cells, cell_end_fill)
        sp, sv = empty_f4(n), empty_f4(n)
        dh, di = to_dev(sorted_hash.astype(np.uint32)), to_dev(index.astype(np.uint32))
        dp, dv = to_dev(pos.astype(np.float32)), to_dev(vel.astype(np.float32))
        L.call("cmag_reorder_find_cell_start", ptr(cs), ptr(ce), ptr(sp), ptr(sv), ptr(dh), ptr(di),
               ptr(dp), ptr(dv), n, num_cells, None)
        out = to_np(cs)[:num_cells], to_np(ce)[:num_cells], to_np(sp)[:n], to_np(sv)[:n]
        del dh, di, dp, dv
        return out

    @staticmethod
    def density(p, sorted_pos, cs, ce, measures_in=None):
        n = sorted_pos.shape[0]
        m = to_dev(measures_in.astype(np.float32)) if measures_in is not None else empty_f4(n)
        dsp, dcs, dce = to_dev(sorted_pos.astype(np.float32)), to_dev(cs), to_dev(ce)
        L.call("cmag_density", P(p), ptr(m), ptr(dsp), ptr(dcs), ptr(dce), n, None)
        out = to_np(m)[:n]
        del dsp, dcs, dce
        return out

    @staticmethod
    def force(p, measures, sorted_pos, sorted_vel, index, cs, ce, acc_in=None):
        n = sorted_pos.shape[0]
        a = to_dev(acc_in.astype(np.float32)) if acc_in is not None else empty_f4(n)
        dm, dsp = to_dev(measures.astype(np.float32)), to_dev(sorted_pos.astype(np.float32))
        dsv, di = to_dev(sorted_vel.astype(np.float32)), to_dev(index.astype(np.uint32))
        dcs, dce = to_dev(cs), to_dev(ce)
        L.call("cmag_force", P(p), ptr(a), ptr(dm), ptr(dsp), ptr(dsv), ptr(di), ptr(dcs), ptr(dce), n, None)
        out = to_np(a)[:n]
        del dm, dsp, dsv, di, dcs, dce
        return out

    @staticmethod
    def pressure_force(p, measures, sorted_pos, index, cs, ce):
        n = sorted_pos.shape[0]
        a = empty_f4(n)
        dm, dsp, di = to_dev(measures.astype(np.float32)), to_dev(sorted_pos.astype(np.float32)), \
            to_dev(index.astype(np.uint32))
        dcs, dce = to_dev(cs), to_dev(ce)
        L.call("cmag_pressure_force", P(p), ptr(a), ptr(dm), ptr(dsp), ptr(di), ptr(dcs), ptr(dce), n, None)
        out = to_np(a)[:n]
        del dm, dsp, di, dcs, dce
        return out

    @staticmethod
    def viscous_force(p, measures, sorted_pos, sorted_vel, index, cs, ce):
        n = sorted_pos.shape[0]
        a = empty_f4(n)
        dm, dsp = to_dev(measures.astype(np.float32)), to_dev(sorted_pos.astype(np.float32))
        dsv, di = to_dev(sorted_vel.astype(np.float32)), to_dev(index.astype(np.uint32))
        dcs, dce = to_dev(cs), to_dev(ce)
        L.call("cmag_viscous_force", P(p), ptr(a), ptr(dm), ptr(dsp), ptr(dsv), ptr(di), ptr(dcs), ptr(dce),
               n, None)
        out = to_np(a)[:n]
        del dm, dsp, dsv, di, dcs, dce
        return out

    @staticmethod
    def integrate(p, pos, vel, vlf, acc, pres=None, wave_time=0.0):
        n = pos.shape[0]
        dp, dv, dl = to_dev(pos.astype(np.float32)), to_dev(vel.astype(np.float32)), to_dev(vlf.astype(np.float32))
        da = to_dev(acc.astype(np.float32))
        dpr = to_dev(pres.astype(np.float32)) if pres is not None else None
        L.call("cmag_integrate", P(p), ptr(dp), ptr(dv), ptr(dl), ptr(da), ptr(dpr) if dpr is not None else None,
               C.c_float(wave_time), n, None)
        out = to_np(dp)[:n], to_np(dv)[:n], to_np(dl)[:n]
        del da, dpr
        return out

    @staticmethod
    def pos_kernel(name, p, pos, *extra):
        n = pos.shape[0]
        d = to_dev(pos.astype(np.float32))
        L.call(name, P(p), ptr(d), *extra, n, None)
        return to_np(d)[:n]


def normwise(got: np.ndarray, want: np.ndarray) -> float:
    """max |got - want| / max |want| over the field"""
    scale = float(np.max(np.abs(want)))
    if scale == 0.0:
        return float(np.max(np.abs(got)))
    return float(np.max(np.abs(got.astype(np.float64) - want.astype(np.float64))) / scale)


class AdjointStages:
    """the adjoint stage launchers (cmag_*_vjp) over torch buffers"""

    @staticmethod
    def integrate_vjp(p, pos, gpos, gvel, gvlf):
        n = pos.shape[0]
        dp = to_dev(pos.astype(np.float32))
        dgp, dgv, dgl = (to_dev(a.astype(np.float32)) for a in (gpos, gvel, gvlf))
        dga = empty_f4(n)
        dgg = torch.zeros((max(n, 1), 4), dtype=torch.float64, device="cuda")
        L.call("cmag_integrate_vjp", P(p), ptr(dp), ptr(dgp), ptr(dgv), ptr(dgl), ptr(dga), ptr(dgg), n, None)
        out = tuple(to_np(t)[:n] for t in (dgp, dgv, dgl, dga, dgg))
        del dp
        return out

    @staticmethod
    def force_vjp(p, gacc_sorted, measures, sorted_pos, sorted_vel, cs, ce):
        n = sorted_pos.shape[0]
        gp, gv = empty_f4(n), empty_f4(n)
        gr = torch.zeros(max(n, 1), dtype=torch.float32, device="cuda")
        gm = torch.zeros(max(n, 1), dtype=torch.float32, device="cuda")
        dga, dm = to_dev(gacc_sorted.astype(np.float32)), to_dev(measures.astype(np.float32))
        dsp, dsv = to_dev(sorted_pos.astype(np.float32)), to_dev(sorted_vel.astype(np.float32))
        dcs, dce = to_dev(cs), to_dev(ce)
        L.call("cmag_force_vjp", P(p), ptr(gp), ptr(gv), ptr(gr), ptr(gm), ptr(dga), ptr(dm), ptr(dsp), ptr(dsv),
               ptr(dcs), ptr(dce), n, None)
        out = to_np(gp)[:n], to_np(gv)[:n], to_np(gr)[:n], to_np(gm)[:n]
        del dga, dm, dsp, dsv, dcs, dce
        return out

    @staticmethod
    def density_vjp(p, gpos_sorted, grho, sorted_pos, cs, ce):
        n = sorted_pos.shape[0]
        gp = to_dev(gpos_sorted.astype(np.float32))
        dgr, dsp = to_dev(grho.astype(np.float32)), to_dev(sorted_pos.astype(np.float32))
        dcs, dce = to_dev(cs), to_dev(ce)
        L.call("cmag_density_vjp", P(p), ptr(gp), ptr(dgr), ptr(dsp), ptr(dcs), ptr(dce), n, None)
        out = to_np(gp)[:n]
        del dgr, dsp, dcs, dce
        return out
