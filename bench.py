"""Benchmark: simulated days per wall hour of the Newton-Krylov time step on
config 1 (256x256x64 = 4.2M cells, dx=dy=5 m, dz=0.25 m, GRF lnKs, four
layers, adaptive steps from dt = 600 s), one GPU, mixed path vs plain FP64.

A "step" is one accepted implicit time step (calc_numsol plus the time-step
control).  Two runs over the same simulated interval:

* plain FP64: the reference's algorithm (D-ILU(0) BiCGStab, the reference's
  FD Jacobian, the reference's dt heuristic) -- W untimed + K timed steps;
* mixed: FP32 vertical-line preconditioner in FP64 BiCGStab with the
  Jacobian made consistent with the residual, replaying the plain run's step
  times, W untimed + K timed steps.

Each timed region is bracketed by device synchronisation.  Inputs are
synthetic (seeded), resident in HBM before the timed region.  `value` is the
mixed path's sim-days/hour; the plain path's figure, the speedup and the
max head difference between the two runs at the last step are reported
beside it.  `roofline` covers the SpMV kernel (algorithmic bytes / average
duration from HIP events on the library's stream); `cpu_baseline` times the
reference build on the host cores on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
GROUPS = ("vg_pair", "assemble", "vg", "residual", "ilu_factor", "ilu_fwd", "ilu_bwd",
          "line_factor", "line_apply", "amg_setup", "amg_numeric", "amg_vcycle", "resid", "spmv")


def spmv_bytes(n, nconn, s_a=8, s_v=8):
    """Algorithmic bytes of one SpMV: (N + M) matrix values + x read + y written."""
    return (n + nconn) * s_a + 2 * n * s_v


def cpu_baseline(p, newton_iters, sim_seconds, sample_newton=6):
    """Time the reference build (oracle/_ref/ref_harness: the reference's own
    modules compiled with amdflang -O2 -fopenmp) on this host's cores: the
    first step's calc_numsol capped at `sample_newton` Newton iterations,
    input reading and set-up excluded.  Its per-iteration cost is scaled by
    the Newton iterations the plain FP64 run (the same algorithm) needed for
    the simulated interval.  Without the reference build, the oracle's C
    restatement is timed on one core instead."""
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if os.path.exists(exe):
        return _reference_baseline(exe, p, newton_iters, sim_seconds, sample_newton)
    return _port_baseline(p, newton_iters, sim_seconds)


def _reference_baseline(exe, p, newton_iters, sim_seconds, sample_newton):
    import subprocess
    import tempfile
    from oracle.deck import write_deck
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    with tempfile.TemporaryDirectory(prefix="msg123_ref_") as d:
        write_deck(p, d)
        subprocess.run([exe], cwd=d, check=True, timeout=900, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL,
                       env=dict(os.environ, OMP_NUM_THREADS=str(threads),
                                MSG123_TIME_NEWTON=str(sample_newton)))
        wall, out_iter = open(os.path.join(d, "timing.txt")).read().split()[:2]
    wall, out_iter = float(wall), int(out_iter)
    per_iter = wall / max(out_iter, 1)
    est = per_iter * newton_iters
    return dict(value=(sim_seconds / 86400.0) / (est / 3600.0), unit="sim-days/hour",
                cores=threads, kind="reference",
                sample=(f"reference build (amdflang -O2 -fopenmp, OMP_NUM_THREADS={threads}) on the "
                        f"same {p.nx}x{p.ny}x{p.nz} deck: calc_numsol of step 1 capped at "
                        f"{out_iter} Newton iterations took {wall:.2f} s ({per_iter:.3f} s/iteration),"
                        f" scaled to the {newton_iters} Newton iterations the plain FP64 run "
                        f"needed"))


def _port_baseline(p, newton_iters, sim_seconds, sample_newton=2):
    import ctypes as C
    from oracle import orc
    om = orc.OracleMesh(p)
    st = orc.OracleState(om).first_step()
    cfg = orc.Settings(sample_newton, p.maxinn_iter, p.criteria, p.init_step, p.max_step,
                       p.end_time)
    t0 = time.perf_counter()
    orc.lib().orc_calc_numsol(C.byref(om.m), C.byref(st.s), C.byref(cfg))
    wall = time.perf_counter() - t0
    per_iter = wall / max(st.s.out_iter, 1)
    est = per_iter * newton_iters
    value = (sim_seconds / 86400.0) / (est / 3600.0)
    return dict(value=value, unit="sim-days/hour", cores=1, kind="port",
                sample=(f"oracle (C restatement of the reference, 1 thread) on the same "
                        f"{p.nx}x{p.ny}x{p.nz} grid: {st.s.out_iter} Newton iterations of step 1 "
                        f"in {wall:.1f} s ({per_iter:.2f} s/iteration), scaled to the "
                        f"{newton_iters} Newton iterations the plain FP64 run needed"))


def timed_run(ctx, warmup, steps, schedule=None):
    """W untimed steps, then K timed steps; returns the timed run, its wall
    time, per-kernel-group times and the warm-up step times."""
    warm_times = np.zeros(0)
    t0 = time.perf_counter()
    if warmup:
        warm_times = ctx.run(nstep_max=warmup, keep_heads=False, reset=True,
                             schedule=schedule)["times"]
    ctx.warmup_wall = time.perf_counter() - t0
    ctx.set_timing(True)
    t0 = time.perf_counter()
    res = ctx.run(nstep_max=steps, keep_heads=False, reset=(warmup == 0), schedule=schedule)
    wall = time.perf_counter() - t0   # run() synchronises the stream before returning
    kern = {g: ctx.timing(g) for g in GROUPS}
    ctx.set_timing(False)
    return res, wall, kern, warm_times


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=46)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="config1")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--plain-only", action="store_true")
    ap.add_argument("--no-amg", action="store_true", help="skip the FP32-AMG path")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain 512 MiB device copy (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs one domain on one GPU; --gpus must be 1"}))
        sys.exit(2)
    if args.calibrate_copy:
        import torch
        n = 64 * 1024 * 1024
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "copies": 3}))
        return

    from msg123_amd import configs, gpu
    nsteps = args.warmup + args.steps
    p = getattr(configs, args.config)(nsteps=nsteps) if args.config == "config1" \
        else getattr(configs, args.config)()
    box = gpu.BoxMesh(p)

    # plain FP64 (the reference's algorithm)
    plain = gpu.Context(box, precon=gpu.PRECON_DILU)
    n, nconn = plain.n, plain.nconn
    res_p, wall_p, kern_p, warm_p = timed_run(plain, args.warmup, args.steps)
    sim_s = float(res_p["dts"].sum())
    newton_p = int(res_p["newton"].sum())
    value_p = (sim_s / 86400.0) / (wall_p / 3600.0)
    spmv_ms, spmv_n = kern_p["spmv"]
    spmv_avg_ms = spmv_ms / max(spmv_n, 1)
    achieved = spmv_bytes(n, nconn) / (spmv_avg_ms * 1e-3) / 1e9
    h_plain = plain.get_state()["head_new"]
    # the step times the plain run took (warm-up + timed), replayed by the mixed run
    sched = np.concatenate([warm_p, res_p["times"]]).astype(np.float32)
    plain.close()

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "spmv_traffic_c1.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get("bytes_per_launch")
    out = {
        "metric": "sim-days/hour at 4.2M cells (config 1), mixed path vs plain FP64, 1 GPU",
        "value": value_p, "unit": "sim-days/hour", "n_gpus": 1, "steps": int(res_p["nstep"]),
        "warmup": args.warmup, "ms_per_step": 1e3 * wall_p / max(res_p["nstep"], 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded GRF soil, config 1 statistics)",
        "config": {"workload": f"{args.config}: {p.nx}x{p.ny}x{p.nz} cells, {args.warmup}+"
                   f"{args.steps} accepted time steps", "cells": n, "sim_seconds": sim_s,
                   "plain": {"sim_days_per_hour": value_p, "wall_s": wall_p,
                             "newton_iterations": newton_p,
                             "bicgstab_iterations": int(res_p["inner"].sum()),
                             "cell_newton_iterations_per_s": n * newton_p / wall_p,
                             "ms_per_step": 1e3 * wall_p / max(res_p["nstep"], 1)}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_spmv_dot (SpMV fused with one dot, FP64, plain run)",
                     "bytes_per_launch": spmv_bytes(n, nconn), "avg_ms": spmv_avg_ms,
                     "launches": spmv_n},
        "kernel_ms": {"plain": {g: round(v[0], 3) for g, v in kern_p.items() if v[1]}},
    }
    if not args.plain_only:
        fast = gpu.Context(box, precon=gpu.PRECON_LINE32, consistent_jacobian=True)
        res_f, wall_f, kern_f, _ = timed_run(fast, args.warmup, args.steps, schedule=sched)
        h_fast = fast.get_state()["head_new"]
        fast.close()
        sim_f = float(res_f["dts"].sum())
        value_f = (sim_f / 86400.0) / (wall_f / 3600.0)
        newton_f = int(res_f["newton"].sum())
        same = bool(np.array_equal(res_f["times"], res_p["times"]))
        out["value"] = value_f
        out["ms_per_step"] = 1e3 * wall_f / max(res_f["nstep"], 1)
        out["dtype"] = "f64 (FP32 preconditioner)"
        out["config"]["mixed"] = {
            "sim_days_per_hour": value_f, "wall_s": wall_f, "newton_iterations": newton_f,
            "bicgstab_iterations": int(res_f["inner"].sum()),
            "cell_newton_iterations_per_s": n * newton_f / wall_f,
            "ms_per_step": 1e3 * wall_f / max(res_f["nstep"], 1),
            "same_step_times_as_plain": same,
            "max_abs_head_diff_vs_plain_m": float(np.abs(h_fast - h_plain).max()
                                                  * box.b.mesh.len_scal),
            "speedup_vs_plain": value_f / value_p}
        out["kernel_ms"]["mixed"] = {g: round(v[0], 3) for g, v in kern_f.items() if v[1]}
    if not args.plain_only and not args.no_amg:
        amg = gpu.Context(box, precon=gpu.PRECON_AMG32, consistent_jacobian=True)
        res_a, wall_a, kern_a, _ = timed_run(amg, args.warmup, args.steps, schedule=sched)
        h_amg = amg.get_state()["head_new"]
        amg_setups, amg_fallbacks = amg.amg_stats()
        value_a = (float(res_a["dts"].sum()) / 86400.0) / (wall_a / 3600.0)
        newton_a = int(res_a["newton"].sum())
        out["config"]["mixed_amg"] = {
            "sim_days_per_hour": value_a, "wall_s": wall_a, "newton_iterations": newton_a,
            "bicgstab_iterations": int(res_a["inner"].sum()),
            "amg_setups": amg_setups, "solves_falling_back_to_line32": amg_fallbacks,
            "warmup_wall_s_incl_hierarchy_setup": amg.warmup_wall,
            "cell_newton_iterations_per_s": n * newton_a / wall_a,
            "same_step_times_as_plain": bool(np.array_equal(res_a["times"], res_p["times"])),
            "max_abs_head_diff_vs_plain_m": float(np.abs(h_amg - h_plain).max()
                                                  * box.b.mesh.len_scal),
            "speedup_vs_plain": value_a / value_p}
        amg.close()
        out["kernel_ms"]["mixed_amg"] = {g: round(v[0], 3) for g, v in kern_a.items() if v[1]}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(p, newton_p, sim_s)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
