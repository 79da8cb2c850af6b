"""Benchmark: simulated days per wall hour of the Newton-Krylov time step on
config 1 (256x256x64 = 4.2M cells, dx=dy=5 m, dz=0.25 m, GRF lnKs, four
layers, 48 adaptive steps from dt = 600 s growing x1.2 up to 3600 s), one GPU.

A "step" is one accepted implicit time step (calc_numsol plus the reference's
time-step control).  W steps run untimed, then exactly K steps are timed with
the device synchronised on both sides.  Inputs are synthetic (seeded) and are
resident in HBM when the timed region starts.

Prints one JSON line with the metric, a `roofline` object for the SpMV kernel
(algorithmic bytes / average kernel duration from HIP events on the library's
stream) and a `cpu_baseline` object (the oracle, one core, on a bounded
sample of the same workload).
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


def spmv_bytes(n, nconn, s_a=8, s_v=8):
    """Algorithmic bytes of one SpMV: (N + M) matrix values + x read + y written."""
    return (n + nconn) * s_a + 2 * n * s_v


def cpu_baseline(p, newton_iters_gpu, sim_seconds, sample_newton=2):
    """Time the oracle (single-threaded C restatement) on the same grid for the
    first `sample_newton` Newton iterations of step 1, then scale its
    per-iteration cost by the Newton iterations the timed GPU steps needed."""
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
    est = per_iter * newton_iters_gpu
    value = (sim_seconds / 86400.0) / (est / 3600.0)
    return dict(value=value, unit="sim-days/hour", cores=1, kind="port",
                sample=(f"oracle (C restatement, 1 thread) on the same 256x256x64 grid: "
                        f"{st.s.out_iter} Newton iterations of step 1 in {wall:.1f} s "
                        f"({per_iter:.2f} s/iteration), scaled to the {newton_iters_gpu} "
                        f"Newton iterations of the timed steps"))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=46)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="config1")
    ap.add_argument("--precon", default="dilu", choices=["dilu", "mc32"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
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
    p = getattr(configs, args.config)(nsteps=args.warmup + args.steps) \
        if args.config == "config1" else getattr(configs, args.config)()
    t_setup = time.perf_counter()
    box = gpu.BoxMesh(p)
    ctx = gpu.Context(box, precon=gpu.PRECON_MC32 if args.precon == "mc32" else gpu.PRECON_DILU)
    n, nconn = ctx.n, ctx.nconn
    setup_s = time.perf_counter() - t_setup

    if args.warmup:
        ctx.run(nstep_max=args.warmup, keep_heads=False, reset=True)
    ctx.set_timing(True)
    t0 = time.perf_counter()
    res = ctx.run(nstep_max=args.steps, keep_heads=False, reset=(args.warmup == 0))
    wall = time.perf_counter() - t0   # run() synchronises the stream before returning
    k = res["nstep"]
    spmv_ms, spmv_n = ctx.timing("spmv")
    groups = {g: ctx.timing(g) for g in ("vg_pair", "assemble", "vg", "residual", "ilu_factor",
                                         "ilu_fwd", "ilu_bwd", "resid", "spmv")}
    ctx.set_timing(False)

    sim_s = float(res["dts"].sum())
    newton = int(res["newton"].sum())
    value = (sim_s / 86400.0) / (wall / 3600.0)
    spmv_avg_ms = spmv_ms / max(spmv_n, 1)
    achieved = spmv_bytes(n, nconn) / (spmv_avg_ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "spmv_traffic_c1.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get("bytes_per_launch")
    out = {
        "metric": "sim-days/hour (config 1, 4.2M cells, plain FP64 D-ILU BiCGStab)"
        if args.precon == "dilu" else "sim-days/hour (config 1, 4.2M cells, FP32-preconditioned)",
        "value": value, "unit": "sim-days/hour", "n_gpus": 1, "steps": k,
        "warmup": args.warmup, "ms_per_step": 1e3 * wall / max(k, 1), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{args.config}: {p.nx}x{p.ny}x{p.nz} cells, "
                   f"{args.warmup}+{args.steps} adaptive steps, precon {args.precon}",
                   "cells": n, "newton_iterations": newton,
                   "sim_seconds": sim_s,
                   "cell_newton_iterations_per_s": n * newton / wall,
                   "setup_s": setup_s},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_spmv_dot (fused SpMV + dot)",
                     "bytes_per_launch": spmv_bytes(n, nconn), "avg_ms": spmv_avg_ms,
                     "launches": spmv_n},
        "kernel_ms": {g: {"total_ms": v[0], "launches": v[1]} for g, v in groups.items()},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(p, newton, sim_s)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
