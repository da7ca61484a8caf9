"""The sub-cell walk's trimming rule, checked by brute force on the CPU.

The DamBreak neighbour passes can walk each cell's fluid by eighth of the
cell along x and, for a target in eighth q, skip the fluid of eighths < q in
the stencil's first cell along x and of eighths > q in its last cell.  That
is only sound if every skipped candidate lies at least 2h away along x.  The
eighth is computed as the GPU computes it (FP32 quotient by the cell size,
then the fraction above its floor, times 8, capped at 7); this test draws
targets and candidates over several cells, including points on eighth and
cell boundaries, and checks that no skipped candidate is closer than 2h (up
to FP32 rounding of the positions)."""
import numpy as np
import pytest

F = np.float32


SUBS = 8  # x_sub in common.hpp


def eighth(x, origin, cs):
    u = (F(x) - F(origin)) / F(cs)
    return np.minimum(SUBS - 1, ((u - np.floor(u)) * F(SUBS)).astype(np.int64)), np.floor(u).astype(np.int64)


@pytest.mark.parametrize("cellcount,ratio", [(1, 1.0), (1, 0.93), (3, 1.0), (2, 0.8)])
def test_skipped_candidates_lie_beyond_2h(cellcount, ratio):
    """ratio = 2h / (cellcount * cell size) <= 1 (the stencil reaches 2h)"""
    rng = np.random.default_rng(100 * cellcount + int(ratio * 100))
    cs = F(0.0052)
    origin = F(-0.0311)
    two_h = F(ratio * cellcount * float(cs))
    n = 200_000
    # targets and candidates over 12 cells, a third of them on eighth boundaries
    cells = rng.integers(0, 12, n)
    frac = rng.uniform(0, 1, n)
    frac[: n // 3] = rng.integers(0, SUBS, n // 3) / SUBS
    xt = (origin + (cells + frac) * cs).astype(F)
    cand_cells = cells + rng.choice([-cellcount, cellcount], n)
    fc = rng.uniform(0, 1, n)
    fc[n // 2: n // 2 + n // 3] = rng.integers(0, SUBS, n // 3) / SUBS
    xc = (origin + (cand_cells + fc) * cs).astype(F)
    qt, ct = eighth(xt, origin, cs)
    qc, cc = eighth(xc, origin, cs)
    lo = cc == ct - cellcount
    hi = cc == ct + cellcount
    skipped = (lo & (qc < qt)) | (hi & (qc > qt))
    assert skipped.sum() > n // 10
    dist = np.abs(xt.astype(np.float64) - xc.astype(np.float64))
    # FP32 positions: a few ulps of the coordinates below 2h at most
    slack = 4 * np.spacing(np.float32(np.abs(xt).max() + np.abs(xc).max()))
    assert np.all(dist[skipped] >= float(two_h) - slack), dist[skipped].min()


def test_rule_is_tight():
    """kept candidates do reach below 2h: the rule does not keep only far ones"""
    rng = np.random.default_rng(3)
    cs, origin = F(0.0052), F(0.0)
    n = 100_000
    xt = (origin + (5 + rng.uniform(0, 1, n)) * cs).astype(F)
    xc = (origin + (4 + rng.uniform(0, 1, n)) * cs).astype(F)
    qt, _ = eighth(xt, origin, cs)
    qc, _ = eighth(xc, origin, cs)
    kept = qc >= qt
    dist = np.abs(xt.astype(np.float64) - xc.astype(np.float64))
    assert np.mean(dist[kept] < float(cs)) > 0.5
