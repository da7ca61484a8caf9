"""Benchmark of the WCSPH timestep on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--dx D] [--hash morton|linear]

Workload: the 3-D dam break at 16,484,816 fluid particles (0.4 x 0.8 x 0.4
column, dx = 0.4/202, in a 1.6 x 1.0 x 0.4 tank with three wall layers on the
floor and four sides, 3,586,416 wall particles), Z-order cell hash over 16x1x1
cell bricks, cell = 2h,
FP32, synthetic initial state.  One step = one update(): hash, stable radix
sort, reorder + cellStart, density (Tait), fused pressure + artificial
viscosity + Lennard-Jones wall force, integrate with clamps.  Inputs are
resident in HBM before the timed region.

Prints one JSON line: particle-steps/s (fluid particles), with
  roofline      the dominant kernel's algorithmic HBM bytes per launch over its
                HIP-event-measured average duration, against 8 TB/s
  batched       256 uneven 2-D channels stepped together (HIP graphs,
                autotuned launch shape and cell size) against the same
                instances as separate System objects stepped in a loop
  adjoint       gradient of a 200-step 16k-particle Poiseuille profile loss
                (forward + hand-written backward) against the plain forward run
  cpu_baseline  the oracle's FP32 update() on this host's cores, on the
                1,024,000-fluid-particle dam break (a bounded sample of steps)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

# algorithmic bytes per particle for each stage (float4 = 16 B, u32 = 4 B);
# every particle (fluid + wall) is charged every stage
ALG_BYTES = {
    "hash": 16 + 8,                      # pos -> hash, index
    "sort": 16,                          # one read + one write of (key, value)
    "reorder": 4 + 4 + 16 + 16 + 32,     # hash, index, gathered pos/velLF -> sortedPos/Vel
    "density": 16 + 16,                  # sortedPos -> measures
    "force": 16 + 16 + 16 + 4 + 16,      # sortedPos/Vel, measures, index -> acc
    # sub-cell walk: rho and p ride in the w of the row copies, no measures read
    "force_sub": 16 + 16 + 4 + 16,
    "integrate": 64 + 48,                # pos, vel, acc, velLF -> pos, vel, velLF
}


# per particle and step for the batched channels (every stage, every particle):
# hash 24, sort 16, reorder 72, density 32, then the fused force pass (sortedPos,
# sortedVel, measures, index -> acc, or -> viscous + pressure force) and
# integrate (pos, vel, acc, velLF [, presF] -> pos, vel, velLF)
BATCH_BYTES = {
    "poiseuille": 24 + 16 + 72 + 32 + (16 + 16 + 16 + 4 + 16) + (64 + 48),
    "peristalsis": 24 + 16 + 72 + 32 + (16 + 16 + 16 + 4 + 32) + (80 + 48),
}


def batch_alg_bytes(scen) -> int:
    """algorithmic bytes of one batched step: the per-particle bytes above and
    each instance's cellStart fill (4 B a cell); the cell-range reads of the
    occupied cells are not charged"""
    from cmag_amd import _lib as L
    from cmag_amd import scenarios as S

    total = 0
    for sc in scen:
        kind = "peristalsis" if sc.params.system == L.SYS_PERISTALSIS else "poiseuille"
        total += BATCH_BYTES[kind] * sc.pos.shape[0] + 4 * S.num_cells(sc.params)
    return total


def cell_bytes(stage: str, num_cells: int, occupied: int) -> int:
    """cell-array traffic per launch: the cellStart fill, the start/end writes of
    occupied cells in reorder, and one start/end read per occupied cell per pass"""
    if stage == "reorder":
        return 4 * num_cells + 8 * occupied
    if stage in ("density", "force"):
        return 8 * occupied
    return 0


def count_pairs(sysobj, L) -> int:
    """in-range neighbour pairs (j != i, |r| < 2h) of the current sorted state"""
    import ctypes as C

    import torch

    n = sysobj.getNumParticles()
    counts = torch.empty(n, dtype=torch.int32, device="cuda")
    p = sysobj.params
    sysobj.synchronize()
    L.call("cmag_count_neighbours", C.byref(p), counts.data_ptr(), sysobj.device_ptr(L.ARRAY_SORTEDPOS),
           sysobj.device_ptr(L.ARRAY_CELLSTART), sysobj.device_ptr(L.ARRAY_CELLEND), n, None)
    torch.cuda.synchronize()
    return int(counts.sum(dtype=torch.int64).item())


def copy_probe(nbytes: int = 1 << 30, reps: int = 5) -> None:
    """plain device copies of known size: the calibration for rocprof's
    FETCH_SIZE / WRITE_SIZE counters (one copy moves nbytes in and out)"""
    import torch

    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()


def measured_traffic(workload_key: str, kernel: str) -> float | None:
    """HBM bytes per launch of kernel from the committed rocprof counter
    summary of this workload, or None"""
    f = ROOT / "profiles" / "traffic.json"
    if not f.exists():
        return None
    try:
        rec = json.loads(f.read_text()).get(workload_key, {})
        return rec.get("kernels", {}).get(kernel, {}).get("hbm_bytes_per_launch")
    except (ValueError, AttributeError):
        return None


# VALU issue peak: one wave64 VALU instruction per CU per cycle (four SIMDs of
# 16 lanes, four cycles per wave64 instruction), 256 CUs at 2.4 GHz
VALU_PEAK_WAVE_INSTS = 256 * 2.4e9


def valu_roofline(workload_key: str, per_stage: dict) -> list | None:
    """Issue-rate roofline of the neighbour passes: the VALU instructions per
    launch from the committed counter summary of this workload over the live
    launch time, against the issue peak"""
    f = ROOT / "profiles" / "valu.json"
    if not f.exists():
        return None
    try:
        rec = json.loads(f.read_text()).get(workload_key, {}).get("kernels", {})
    except ValueError:
        return None
    out = []
    for stage in ("force", "density"):
        k = rec.get(stage)
        if not k or stage not in per_stage or not per_stage[stage]["avg_ms"]:
            continue
        achieved = k["valu_insts_per_launch"] / (per_stage[stage]["avg_ms"] * 1e-3)
        out.append({"bound": "valu", "kernel": stage, "achieved": round(achieved / 1e9, 1),
                    "peak": round(VALU_PEAK_WAVE_INSTS / 1e9, 1), "unit": "G wave64 VALU instr/s",
                    "frac": round(achieved / VALU_PEAK_WAVE_INSTS, 3),
                    "valu_insts_per_wave": k["valu_insts_per_wave"],
                    "lane_utilization": k.get("valu_lane_utilization")})
    return out or None


def bench_batch(steps: int = 1000, plain_steps: int = 20, tune_cells: bool = True) -> dict:
    """The batched workload: 256 independent 2-D channels (half Poiseuille, half
    peristaltic, 1k..64k particles each) stepped together in one launch sequence
    per step (HIP graphs), against the same instances as 256 separate System
    objects stepped one after another (the per-instance plain loop)."""
    import torch

    from cmag_amd import _lib as L
    from cmag_amd import scenarios as S
    from cmag_amd.batch import Batch, autotune_cells
    from cmag_amd.systems import ParamSystem

    def instances(cf):
        return [(sc.params, sc.pos, sc.vel) for sc in S.channel_batch(cell_factor=cf)]

    t_tune = time.perf_counter()
    if tune_cells:
        tuned = autotune_cells(instances, cell_factors=(1.0, 1.3, 2.0), steps=20, steps_per_trial=3)
    else:
        tuned = {"cell_factor": 1.0, "tuning": None, "trials": []}
    t_tune = time.perf_counter() - t_tune
    scen = S.channel_batch(cell_factor=tuned["cell_factor"])
    n_total = sum(sc.pos.shape[0] for sc in scen)
    b = Batch([(sc.params, sc.pos, sc.vel) for sc in scen])
    if tuned["tuning"]:
        b.set_tuning(**tuned["tuning"])
    else:
        b.autotune(3)
    b.run(20)
    b.synchronize()
    t0 = time.perf_counter()
    b.run(steps)
    b.synchronize()
    wall = time.perf_counter() - t0
    b.enable_timing(True)
    b.stage_times()
    b.run(10, use_graph=False)
    st = {k: round(v[0] / max(v[1], 1), 4) for k, v in b.stage_times().items()}
    b.enable_timing(False)
    finite = all(np.isfinite(b.getArray(k, L.ARRAY_POSITION)).all() for k in range(b.m))
    tuning = b.tuning
    b.close()

    systems = [ParamSystem.from_state(sc.params, sc.pos, sc.vel) for sc in scen]
    for s_ in systems:
        s_.update()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(plain_steps):
        for s_ in systems:
            s_.update()
    torch.cuda.synchronize()
    plain_wall = time.perf_counter() - t0
    # the same loop with each System replaying its own one-step graph where it can
    t0 = time.perf_counter()
    for s_ in systems:
        s_.run(plain_steps, use_graph=True)
    torch.cuda.synchronize()
    plain_graph_wall = time.perf_counter() - t0
    for s_ in systems:
        s_.close()
    batched = n_total * steps / wall
    plain = n_total * plain_steps / plain_wall
    plain_graph = n_total * plain_steps / plain_graph_wall
    step_bytes = batch_alg_bytes(scen)
    step_gbs = step_bytes / (wall / steps) / 1e9
    return {
        "workload": (f"256 2-D channels (128 Poiseuille nu~logU(1e-3,1e-1), 128 peristaltic A/H~U(0.1,0.5), "
                     f"wavelength~U(0.5,2)W), {n_total:,} particles, counts log-uniform 1k..64k"),
        "num_particles": n_total,
        "steps": steps,
        "ms_per_step": 1e3 * wall / steps,
        "particle_steps_per_s": batched,
        "stages_ms": st,
        "tuning": tuning,
        "cell_factor": tuned["cell_factor"],
        "autotune_s": round(t_tune, 2),
        "autotune_trials": [{"cell_factor": t["cell_factor"], "ms_per_step": round(t["ms_per_step"], 3)}
                            for t in tuned["trials"]],
        "finite_after": bool(finite),
        "plain_loop_particle_steps_per_s": plain,
        "plain_loop_steps": plain_steps,
        "plain_loop_graph_particle_steps_per_s": plain_graph,
        "speedup_vs_plain_loop": batched / plain,
        "speedup_vs_plain_loop_graphs": batched / plain_graph,
        "roofline": {"bound": "hbm", "kernel": "whole step", "achieved": round(step_gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(step_gbs / HBM_PEAK_GBS, 4), "alg_bytes_per_step": step_bytes,
                     "traffic": None},
    }


def bench_adjoint(reps: int = 3) -> dict:
    """The differentiable workload: gradient of the 64-bin x-velocity profile
    loss of a 200-step, 16,384-fluid-particle Poiseuille run with respect to
    viscosity and body force (checkpoints every 20 steps), timed forward +
    backward against the plain 200-step forward run of the same system."""
    import torch

    from cmag_amd import scenarios as S
    from cmag_amd.adjoint import ChannelRollout, profile, profile_loss
    from cmag_amd.systems import ParamSystem

    d = S.adjoint_channel()
    sc, P0, Ps = d["scenario"], d["params0"], d["params_star"]
    steps, bins = d["steps"], d["bins"]
    n = sc.pos.shape[0]
    # target profile from (nu*, g*) on the GPU's own forward run
    tgt_run = ChannelRollout(Ps, sc.pos, sc.vel, steps=steps, checkpoint_every=d["checkpoint_every"])
    with torch.no_grad():
        mu_s = torch.tensor(float(Ps.mu), dtype=torch.float64)
        g_s = torch.tensor([float(Ps.gravity[a]) for a in range(3)], dtype=torch.float64)
        pos_t, vel_t = tgt_run(mu_s, g_s)
    target = np.nan_to_num(profile(pos_t, vel_t, P0, bins))
    run = ChannelRollout(P0, sc.pos, sc.vel, steps=steps, checkpoint_every=d["checkpoint_every"])

    def grad_once():
        mu = torch.tensor(float(P0.mu), dtype=torch.float64, requires_grad=True)
        g = torch.tensor([float(P0.gravity[a]) for a in range(3)], dtype=torch.float64, requires_grad=True)
        pos, vel = run(mu, g)
        loss = profile_loss(pos, vel, P0, target, bins)
        loss.backward()
        torch.cuda.synchronize()
        return float(loss.detach()), float(mu.grad), float(g.grad[0])

    grad_once()
    t0 = time.perf_counter()
    for _ in range(reps):
        loss, dmu, dg = grad_once()
    t_grad = (time.perf_counter() - t0) / reps

    plain = ParamSystem.from_state(P0, sc.pos, sc.vel)
    plain.run(steps, use_graph=True)
    plain.synchronize()
    plain = ParamSystem.from_state(P0, sc.pos, sc.vel)
    plain.synchronize()
    t0 = time.perf_counter()
    plain.run(steps, use_graph=True)
    plain.synchronize()
    t_plain = time.perf_counter() - t0
    plain.close()
    return {
        "workload": (f"2-D Poiseuille, {sc.num_fluid:,} fluid + {n - sc.num_fluid:,} wall particles, {steps} steps, "
                     f"checkpoint every {d['checkpoint_every']}, loss over {bins} y-bins, target at "
                     f"nu*={d['nu_star']}, gradient at nu0={d['nu0']}"),
        "loss": loss,
        "dL_dnu": 1000.0 * dmu,
        "dL_dg": dg,
        "forward_backward_ms": 1e3 * t_grad,
        "plain_forward_ms": 1e3 * t_plain,
        "ratio_vs_plain": t_grad / t_plain,
        "particle_steps_per_s": n * steps / t_grad,
        "plain_particle_steps_per_s": n * steps / t_plain,
    }


def cpu_baseline(target_seconds: float = 12.0) -> dict:
    """The oracle's FP32 update() (the reference's loop bodies restated in C,
    OpenMP over particles) on the 1M-fluid dam break, timed on this host.
    Threads: the cores this process may run on, at most 16 -- the host-core
    share one GPU's job gets on the GPU hosts (OMP_NUM_THREADS is 16 there);
    nproc and the affinity count are reported beside it.  The library is the
    AVX-512 build (x86-64-v4) when the host has AVX-512, else the AVX2 one."""
    from cmag_amd import scenarios as S
    from oracle import oracle as O

    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    cores = max(1, min(affinity, 16))
    isa = "x86-64-v3"
    v4 = O.HERE / "liboracle_v4.so"
    if v4.exists() and O.host_has_avx512():
        O.use_default(v4)
        isa = "x86-64-v4"
    O.set_threads(cores)
    sc = S.dambreak3d_particles(1_024_000, hash_order=1)
    sim = O.Simulation(sc.params, sc.pos, sc.vel)
    sim.update(1)  # warm-up
    steps, t0 = 0, time.perf_counter()
    while True:
        sim.update(1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= target_seconds or steps >= 50:
            break
    return {
        "value": sc.num_fluid * steps / el,
        "unit": "particle-steps/s",
        "cores": cores,
        "nproc": os.cpu_count(),
        "affinity_cores": affinity,
        "kind": "port",
        "isa": isa,
        "sample": (f"oracle FP32 update() of the 3-D dam break with {sc.num_fluid:,} fluid + "
                   f"{sc.meta['num_wall']:,} wall particles, {steps} steps after 1 warm-up step, "
                   f"{cores} OpenMP threads, {isa} build"),
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--dx", type=float, default=0.4 / 202)
    ap.add_argument("--hash", choices=["morton", "linear"], default="morton")
    ap.add_argument("--morton-xbits", type=int, default=7,
                    help="Z order over bricks of 2^xbits cells along x (0 = plain Morton)")
    ap.add_argument("--relax", type=float, default=0.05,
                    help="seconds of damped settling behind the gate before the warm-up steps (0: start from the "
                         "lattice)")
    ap.add_argument("--alpha", type=float, default=0.01,
                    help="artificial viscosity of the timed run (the workload names 0.01; the reference hard-codes "
                         "0.38, fluid_kernel.cu:229)")
    ap.add_argument("--relax-eager", action="store_true",
                    help="step the settling phase eagerly rather than by graph replay (for counter profiles)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-channel comparison")
    ap.add_argument("--batch-steps", type=int, default=1000)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable-run comparison")
    ap.add_argument("--workload", choices=["dambreak", "batch", "adjoint"], default="dambreak",
                    help="batch / adjoint: only that comparison (for profiling)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="after the timed steps, run 1 GiB device copies (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); the library steps a system on one GPU (no multi-GPU path)",
              file=sys.stderr)
        return 2

    import torch  # noqa: F401  (initialises the ROCm runtime the library shares)

    if args.workload == "batch":
        res = bench_batch(steps=args.batch_steps)
        print(json.dumps({"batched": res}), flush=True)
        return 0
    if args.workload == "adjoint":
        print(json.dumps({"adjoint": bench_adjoint()}), flush=True)
        return 0

    from cmag_amd import _lib as L
    from cmag_amd import scenarios as S
    from cmag_amd.systems import ParamSystem

    L.lib()
    order = L.HASH_MORTON if args.hash == "morton" else L.HASH_LINEAR
    hash_name = (f"Z-order over {1 << args.morton_xbits}x1x1 cell bricks" if args.hash == "morton"
                 else "linear")
    sc = S.dambreak3d(dx=args.dx, hash_order=order, morton_xbits=args.morton_xbits, alpha=args.alpha)
    n_total = sc.pos.shape[0]
    print(f"# workload: {sc.num_fluid} fluid + {sc.meta['num_wall']} wall particles, grid "
          f"{tuple(sc.params.grid_size)}, hash {args.hash}", file=sys.stderr, flush=True)
    sysobj = ParamSystem.from_state(sc.params, sc.pos, sc.vel)
    relaxed = args.relax > 0 and sc.meta["release_shifts"] > 0
    if relaxed:
        # the column settles behind its gate first: a lattice start has a
        # surface transient of several m/s in its first hundred steps
        t_relax = time.perf_counter()
        last = [time.perf_counter()]

        def report(done, total):
            if time.perf_counter() - last[0] > 20.0:
                last[0] = time.perf_counter()
                print(f"# relaxing: {done} / {total} steps", file=sys.stderr, flush=True)

        S.relax_gated(sysobj, sc, seconds=args.relax, alpha=2.0, zero_every=0.005, free_last=0.005, report=report,
                      use_graph=not args.relax_eager)
        sysobj.synchronize()
        print(f"# relaxed {args.relax} s behind the gate in {time.perf_counter() - t_relax:.1f} s",
              file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        sysobj.update()
    sysobj.synchronize()
    cs = sysobj.getUintArray(L.ARRAY_CELLSTART)
    occupied = int(np.count_nonzero(cs != L.CELL_EMPTY))
    num_cells = cs.size
    del cs
    pairs = count_pairs(sysobj, L)

    sysobj.enable_timing(True)
    sysobj.stage_times()
    sysobj.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sysobj.update()
    sysobj.synchronize()
    t1 = time.perf_counter()
    stages = sysobj.stage_times()
    sysobj.enable_timing(False)
    if args.copy_probe:
        copy_probe()
    wall_s = t1 - t0
    ms_per_step = 1000.0 * wall_s / args.steps
    speed = np.linalg.norm(sysobj.getArray(L.ARRAY_VELOCITY, 0, sc.num_fluid)[:, :3], axis=1)
    state_after = {"t_s": round(sysobj.getElapsedTime(), 5), "vmax": round(float(speed.max()), 3),
                   "v99": round(float(np.quantile(speed, 0.99)), 3), "v50": round(float(np.median(speed)), 4)}
    del speed

    per_stage = {}
    for name, (ms, cnt) in stages.items():
        avg_ms = ms / max(cnt, 1)
        key = "force_sub" if name == "force" and sysobj.subcell_walk else name
        alg = ALG_BYTES[key] * n_total + cell_bytes(name, num_cells, occupied)
        per_stage[name] = {"avg_ms": round(avg_ms, 4), "alg_bytes": alg,
                           "alg_GBs": round(alg / (avg_ms * 1e-3) / 1e9, 1) if avg_ms > 0 else None}
    dominant = max(per_stage, key=lambda k: per_stage[k]["avg_ms"])
    d = per_stage[dominant]
    achieved = d["alg_bytes"] / (d["avg_ms"] * 1e-3) / 1e9
    step_bytes = sum(v["alg_bytes"] for v in per_stage.values())
    print(f"# stages: {json.dumps(per_stage)}", file=sys.stderr, flush=True)
    workload_key = f"dambreak3d_dx{args.dx:.6g}_{args.hash}{args.morton_xbits if args.hash == 'morton' else ''}"
    traffic = measured_traffic(workload_key, dominant)

    result = {
        "metric": "particle-steps/s + HBM GB/s at 16M particles; batched & adjoint vs plain",
        "value": sc.num_fluid * args.steps / wall_s,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"3-D dam break, {sc.num_fluid:,} fluid + {sc.meta['num_wall']:,} wall particles, "
                         f"dx={args.dx:.6g}, h=1.3dx, cell=2h, {hash_name} cell hash, Tait, alpha={args.alpha:g}; "
                         + (f"the column behind its gate after {args.relax} s of damped settling"
                            if relaxed else "the column from rest behind its gate")),
            # fluid speeds (m/s) after the timed steps
            "fluid_state_after": state_after,
            "num_fluid": sc.num_fluid,
            "num_total": n_total,
            "grid": list(sc.params.grid_size),
            "occupied_cells": occupied,
            "step_alg_bytes": step_bytes,
            "step_alg_GBs": round(step_bytes * args.steps / wall_s / 1e9, 1),
            "stages_ms": {k: v["avg_ms"] for k, v in per_stage.items()},
            "neighbour_pairs_per_pass": pairs,
            # in-range ordered pairs (j != i) of the density and force passes per second
            "neighbour_pairs_per_s": 2.0 * pairs * args.steps / wall_s,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dominant,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
        },
    }
    valu = valu_roofline(workload_key, per_stage)
    if valu:
        result["valu_roofline"] = valu
    if not args.no_batch:
        result["batched"] = bench_batch(steps=args.batch_steps)
    if not args.no_adjoint:
        result["adjoint"] = bench_adjoint()
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline()
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
