"""Benchmark: simulated days per wall hour of the Newton-Krylov time step on
config 1 (256x256x64 = 4.2M cells, dx=dy=5 m, dz=0.25 m, GRF lnKs, four
layers, adaptive steps from dt = 600 s), one GPU, mixed + learned path vs
plain FP64.

A "step" is one accepted implicit time step (calc_numsol plus the time-step
control).  Every path covers the same simulated interval, the reference
build's own step times (tests/golden/ref_c1.npz: the reference's 48 accepted
steps of config 1, its first step cut from 600 s to 18.75 s), and every
path's heads at the fixture's 4095 sample cells are compared with the
reference's after every step (`heads_vs_reference`):

* plain FP64: the reference's algorithm (D-ILU(0) BiCGStab, the reference's
  FD Jacobian, Eisenstat-Walker forcing and line search);
* mixed: FP32 vertical-line preconditioner inside FP64 BiCGStab, the
  Jacobian consistent with the residual, the constitutive law in its
  log/exp form, forcing capped at 0.03, the line-search slope from the
  Krylov residual and no beta-condition refinement of the line search;
* mixed + learned (`value`): the mixed path with the BF16 conv-net warm
  start (msg123_amd/warmstart_c1.npy, trained on the GPU by this library's
  kernels on trajectories of other seeds); its gate tries the network on the
  first two timed steps (solved again from persistence), and uses it only if
  it cost less than 0.8 of persistence's Newton iterations there; the line
  says which path `value` measures (`value_path`);
* mixed AMG: the FP32 smoothed-aggregation V-cycle inside FP64 BiCGStab,
  with the reference's forcing and the mixed path's slope and line search.

W warm-up steps, then the fused kernels' launch shapes are tuned on the
state (msg123_autotune), then K timed steps; each timed region is bracketed
by device synchronisation; inputs are synthetic (seeded) and resident in HBM
before it.  `roofline` covers the SpMV kernel (algorithmic bytes / average
duration from HIP events on the library's stream), `roofline_residual` and
`roofline_assembly` the fused residual and Jacobian kernels;
`cpu_baseline` times the reference build on the host cores on one whole
accepted step of the same workload.

`--config config2` runs the seasonal-recharge case over a bounded span with
the four warm starts (seasonal_bench); `--config config3` the ensemble;
`--config config4` the 134M-cell stress case.
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

from msg123_amd.gpu import AMG_NEWTON, MIXED_NEWTON  # noqa: E402

HBM_PEAK_GBS = 8000.0
GROUPS = ("vg_pair", "assemble", "vg", "residual", "ilu_factor", "ilu_fwd", "ilu_bwd",
          "line_factor", "line_apply", "amg_setup", "amg_numeric", "amg_vcycle", "resid", "spmv")
WEIGHTS = os.path.join(ROOT, "msg123_amd", "warmstart_c1.npy")


def residual_bytes(n, ns):
    """Algorithmic bytes of one residual evaluation (the fused kernel
    k_fused<kResid>, constitutive law on chip): per cell h, h_old, S_old
    (24 B), the five float32 soil parameters (20 B) and the conductances of
    the E, S and lower faces (24 B) read, F and S written (16 B); face
    geometry, volumes and cell tops are per-layer tables; per top cell seven
    surface arrays (56 B)."""
    return 84 * n + 56 * ns


def assembly_bytes(n, ns, s_a=8):
    """Algorithmic bytes of one Jacobian assembly (k_fused<kAssem>): the
    residual's 68 B per cell read, rhs and S written (16 B), the seven
    diagonals written (s_a bytes each: 4 for the FP32 Newton systems of the
    line-preconditioned path); per top cell seven surface arrays."""
    return (84 + 7 * s_a) * n + 56 * ns


def spmv_bytes(n, nconn, s_a=8, s_v=8):
    """Algorithmic bytes of one SpMV: (N + M) matrix values + x read + y written."""
    return (n + nconn) * s_a + 2 * n * s_v


def cpu_baseline(p, dt0):
    """Time the reference build (oracle/_ref/ref_harness: the reference's own
    modules compiled with amdflang -O2 -fopenmp) on this host's cores on one
    whole accepted time step of the workload: the deck with its first step
    set to dt0, the reference's first accepted step (18.75 s on config 1,
    after its 600 s attempt was halved five times), solved by calc_numsol to
    convergence; input reading and set-up are excluded.  sim-days/hour =
    dt0 / wall.  The 1-thread figure times the first two Newton iterations of
    the same step and scales them to the step's count.  Without the
    reference build, the oracle's C restatement is timed on one core."""
    import dataclasses
    q = dataclasses.replace(p, init_step=float(dt0))
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if os.path.exists(exe):
        return _reference_baseline(exe, q)
    return _port_baseline(q)


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def _reference_baseline(exe, p):
    import subprocess
    import tempfile
    from oracle.deck import write_deck
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()

    def timed(nthreads, cap):
        with tempfile.TemporaryDirectory(prefix="msg123_ref_") as d:
            write_deck(p, d)
            subprocess.run([exe], cwd=d, check=True, timeout=900, stdout=subprocess.DEVNULL,
                           stderr=subprocess.DEVNULL,
                           env=dict(os.environ, OMP_NUM_THREADS=str(nthreads),
                                    OMP_PROC_BIND="close", OMP_PLACES="cores",
                                    MSG123_TIME_NEWTON=str(cap)))
            wall, out_iter = open(os.path.join(d, "timing.txt")).read().split()[:2]
        return float(wall), int(out_iter)

    _say(f"cpu baseline: reference build, {threads} threads, one accepted step")
    wall, its = timed(threads, p.maxout_iter)
    _say(f"cpu baseline: {its} Newton iterations in {wall:.1f} s; 1 thread, 2 iterations")
    wall1, its1 = timed(1, 2)
    value = (p.init_step / 86400.0) / (wall / 3600.0)
    value1 = (p.init_step / 86400.0) / (wall1 / its1 * its / 3600.0)
    return dict(value=value, unit="sim-days/hour", cores=threads, kind="reference",
                value_1_thread=value1,
                sample=(f"reference build (amdflang -O2 -fopenmp, OMP_NUM_THREADS={threads}, "
                        f"OMP_PROC_BIND=close, OMP_PLACES=cores) on {_cpu_model()}, the same "
                        f"{p.nx}x{p.ny}x{p.nz} deck: one whole accepted step of "
                        f"{p.init_step:g} s (the workload's first accepted step), "
                        f"{its} Newton iterations in {wall:.2f} s; 1 thread: {wall1 / its1:.2f} s "
                        f"per Newton iteration over the first {its1}, scaled to {its}"))


def _port_baseline(p):
    import ctypes as C
    from oracle import orc
    om = orc.OracleMesh(p)
    st = orc.OracleState(om).first_step()
    cfg = orc.Settings(p.maxout_iter, p.maxinn_iter, p.criteria, p.init_step, p.max_step,
                       p.end_time)
    t0 = time.perf_counter()
    orc.lib().orc_calc_numsol(C.byref(om.m), C.byref(st.s), C.byref(cfg))
    wall = time.perf_counter() - t0
    value = (p.init_step / 86400.0) / (wall / 3600.0)
    return dict(value=value, unit="sim-days/hour", cores=1, kind="port",
                sample=(f"oracle (C restatement of the reference, 1 thread) on the same "
                        f"{p.nx}x{p.ny}x{p.nz} grid: one accepted step of {p.init_step:g} s "
                        f"({st.s.out_iter} Newton iterations) in {wall:.1f} s"))


KERNEL_GROUPS = False   # --kernel-groups: HIP events around every kernel group


def _say(msg):
    """Progress of the run (stderr; the JSON line goes to stdout)."""
    print(msg, file=sys.stderr, flush=True)


def timed_run(ctx, warmup, steps, schedule=None, warm_start=0, probes=None, spmv_events=True):
    """W untimed steps, then K timed steps; returns the timed run, its wall
    time, per-kernel-group times and the warm-up step times.  The mass
    balance of every step is accumulated (one extra residual-sized kernel per
    accepted step, inside the timed region).  Only the SpMV group is timed
    with HIP events (spmv_events; the pair around each launch adds about 2%
    to the wall clock), every group with KERNEL_GROUPS (about 6%).  The warm-up steps start from persistence (the
    reference's start), so a learned start's gate tries the network on the
    first timed steps whatever W is, and its cost is inside the timed
    region."""
    ctx.set_mass_balance(True)
    warm_times = np.zeros(0)
    t0 = time.perf_counter()
    if warmup:
        wr = ctx.run(nstep_max=warmup, keep_heads=False, reset=True, schedule=schedule,
                     probes=probes)
        warm_times = wr["times"]
        ctx.warm_probes = wr["probe_heads"]
    ctx.warmup_wall = time.perf_counter() - t0
    # launch shapes of the fused residual / assembly kernels, timed on this
    # context's state (msg123_autotune; set-up, untimed)
    ctx.tune = ctx.autotune(3) if warmup else None
    ctx.set_timing(True if KERNEL_GROUPS else ("spmv" if spmv_events else False))
    t0 = time.perf_counter()
    res = ctx.run(nstep_max=steps, keep_heads=False, reset=(warmup == 0), schedule=schedule,
                  warm_start=warm_start, probes=probes)
    wall = time.perf_counter() - t0   # run() synchronises the stream before returning
    kern = {g: ctx.timing(g) for g in GROUPS}
    ctx.set_timing(False)
    return res, wall, kern, warm_times


def summary(res, wall, n, ctx):
    sim = float(res["dts"].sum())
    newton = int(res["newton"].sum())
    mb = ctx.mass_balance()
    return {"sim_days_per_hour": (sim / 86400.0) / (wall / 3600.0), "wall_s": wall,
            "newton_iterations": newton, "bicgstab_iterations": int(res["inner"].sum()),
            "cell_newton_iterations_per_s": n * newton / wall,
            "ms_per_step": 1e3 * wall / max(res["nstep"], 1),
            # cumulative |sum F(h^{n+1}) dt| / sum |recharge| dt over all steps
            "mass_balance_rel_error": abs(mb["residual"]) / mb["recharge_abs"],
            "mass_balance_total_rel": abs(mb["total"]) / mb["recharge_abs"]}


def vs_reference(ctx, res, ref, warmup):
    """Heads at the reference fixture's sample cells after every step (warm-up
    and timed) against the reference build's own run of the same steps."""
    got = res["probe_heads"] if not warmup else np.concatenate([ctx.warm_probes,
                                                                 res["probe_heads"]])
    k = got.shape[0]
    d = np.abs(got - ref["step_heads_sample"][:k]) * float(ref["len_scal"])
    return {"steps_compared": int(k), "cells_compared": int(d.shape[1]),
            "max_abs_head_diff_m": float(d.max()),
            "cells_beyond_1e-6_m": int(np.count_nonzero((d > 1e-6).any(axis=0))),
            "max_per_step_m": [float(f"{x:.3g}") for x in d.max(axis=1)]}


def ensemble_bench(args):
    """Config 3: 64 independent boxes of uneven size (13.0M cells in all),
    24 hourly steps each with the reference's time control, mixed path; all
    instances at once (one context and stream per instance, 16 host threads)
    against one after another."""
    from msg123_amd import configs, ensemble
    probs = configs.config3()
    ensemble.run(probs[:4], threads=4)          # warm-up: library and kernels loaded
    seq = ensemble.run(probs, threads=1)
    con = ensemble.run(probs, threads=16)
    same = all(np.array_equal(a["times"], b["times"]) and a["newton"] == b["newton"]
               for a, b in zip(seq["instances"], con["instances"]))
    return {
        "metric": "instance-steps/s, config 3 ensemble (64 boxes, 13.0M cells), 1 GPU",
        "value": con["instance_steps_per_s"], "unit": "instance-steps/s", "n_gpus": 1,
        "steps": con["instance_steps"], "warmup": 0,
        "ms_per_step": 1e3 * con["wall_s"] / max(con["instance_steps"], 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (FP32 preconditioner)", "data": "synthetic (seeded iid soil per instance)",
        "config": {"workload": "config3: 64 boxes nx,ny in 32..128, nz in 16..48, 24 x 3600 s",
                   "cells": con["cells"], "threads": 16,
                   "concurrent": {k: con[k] for k in ("wall_s", "instance_steps_per_s",
                                                       "solve_p50_s", "solve_p99_s")},
                   "sequential": {k: seq[k] for k in ("wall_s", "instance_steps_per_s",
                                                       "solve_p50_s", "solve_p99_s")},
                   "same_step_times_and_newton_counts": bool(same),
                   "slowest_instances_sequential": [
                       {k: r[k] for k in ("name", "cells", "steps", "newton", "solve_s")}
                       for r in sorted(seq["instances"], key=lambda r: -r["solve_s"])[:4]],
                   "newton_iterations": sum(r["newton"] for r in con["instances"])},
    }


def seasonal_bench(args):
    """Config 2 over a bounded span: the 4.2M-cell config-1 grid (seed 2)
    under daily seasonal recharge records, the mixed path.  The span is the
    first `--steps` accepted steps of the reference's time control from
    dt = 600 s (the early steps are the stiff ones: 7 to 20 Newton iterations
    at dt of a few minutes).  Every start then replays those step times to
    the same end time: persistence (the reference's start), the linear
    extrapolation, the better of the two by residual, and the learned start
    (the network trained on config-1 statistics, which config 2 shares; its
    two shadow trials are counted in its wall time).  A step that fails from
    some start is halved and sub-stepped, and those steps count against it.
    Reports Newton iterations per step and wall time per start, and the
    learned start's gain over persistence."""
    from msg123_amd import configs, gpu
    steps = args.steps
    p = configs.config2(ndays=2)
    box = gpu.BoxMesh(p)
    n = p.ncell
    L = box.b.mesh.len_scal

    def mixed_ctx():
        ctx = gpu.Context(box, precon=gpu.PRECON_LINE32, consistent_jacobian=True)
        ctx.set_newton(**MIXED_NEWTON)
        ctx.set_law(gpu.LAW_LOGEXP)
        return ctx

    def chunks(ctx, name, sched, mode, cap):
        newton, dts, done = [], [], 0
        t0 = time.perf_counter()
        while done < cap:
            m = min(25, cap - done)
            r = ctx.run(nstep_max=m, keep_heads=False, reset=(done == 0), schedule=sched,
                        warm_start=mode, max_fail=16)
            newton += r["newton"].tolist()
            dts += r["dts"].tolist()
            done += r["nstep"]
            t_now = float(np.sum(dts))
            print(f"config2 {name}: {done} steps, t = {t_now:.0f} s, "
                  f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
            if r["nstep"] < m:
                break
        return newton, dts, done, time.perf_counter() - t0

    # the span: the reference's time control from persistence
    ctx = mixed_ctx()
    sched = np.zeros(0, np.float32)
    done = 0
    while done < steps:
        r = ctx.run(nstep_max=min(25, steps - done), keep_heads=False, reset=(done == 0))
        sched = np.concatenate([sched, r["times"].astype(np.float32)])
        done += r["nstep"]
        print(f"config2 span: {done} steps, t = {float(sched[-1]):.0f} s", file=sys.stderr,
              flush=True)
        if r["nstep"] == 0:
            break
    ctx.close()
    out, h_base, net = {}, None, None
    for name, mode in (("persist", gpu.WARM_PERSIST), ("linear", gpu.WARM_LINEAR),
                       ("select", gpu.WARM_SELECT), ("learned", gpu.WARM_LEARNED)):
        ctx = mixed_ctx()
        if mode == gpu.WARM_LEARNED:
            net = gpu.Net(p.nx, p.ny, p.nz)
            net.params(np.load(WEIGHTS))
            ctx.set_warm_net(net, np.log(p.ks), float(p.rech_mmh[0]) * 1e-3 / 3600.0)
        newton, dts, k, wall = chunks(ctx, name, sched, mode, 4 * len(sched))
        h = ctx.get_state()["head_new"]
        if h_base is None:
            h_base = h
        sim = float(np.sum(dts))
        out[name] = {"steps": k, "sub_steps": k - len(sched),
                     "newton_iterations": int(np.sum(newton)),
                     "newton_per_scheduled_step": float(np.sum(newton)) / len(sched),
                     "wall_s": wall, "sim_s": sim,
                     "sim_days_per_hour": (sim / 86400.0) / (wall / 3600.0),
                     "max_head_diff_vs_persist_m": float(np.abs(h - h_base).max() * L),
                     "final_head_sum": float(h.sum())}
        if mode != gpu.WARM_PERSIST:
            out[name]["warm_stats"] = ctx.warm_stats()
        ctx.close()
    if net is not None:
        net.close()
    best = max(out, key=lambda k: out[k]["sim_days_per_hour"])
    lr, ps = out["learned"], out["persist"]
    return {
        "metric": "sim-days/hour, config 2 (4.2M cells, seasonal recharge records), "
                  "bounded span, mixed path, 1 GPU",
        "value": out[best]["sim_days_per_hour"], "unit": "sim-days/hour", "n_gpus": 1,
        "steps": len(sched), "warmup": 0,
        "ms_per_step": 1e3 * out[best]["wall_s"] / max(out[best]["steps"], 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (FP32 preconditioner)",
        "data": "synthetic (seeded GRF lnKs, config-1 statistics, seed 2)",
        "config": {"workload": f"config2: 256x256x64 cells, the first {len(sched)} steps of the "
                               f"reference's time control from dt = 600 s "
                               f"({float(sched[-1]):.0f} s simulated), daily records of "
                               f"1e-7 (1 + 0.8 sin(2 pi t / 365 d)) m/s",
                   "cells": n, "value_start": best, "starts": out,
                   "learned_newton_vs_persist": lr["newton_iterations"] /
                   max(ps["newton_iterations"], 1),
                   "learned_wall_vs_persist": lr["wall_s"] / ps["wall_s"],
                   "learned_network_used": bool(lr["warm_stats"]["net_on"] == 1),
                   "learned_start_gain": bool(lr["wall_s"] < ps["wall_s"])},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=46)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="config1")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--plain-only", action="store_true")
    ap.add_argument("--no-amg", action="store_true", help="skip the FP32-AMG path")
    ap.add_argument("--control", default="host", choices=("host", "device", "graph"),
                    help="who drives the mixed paths' Newton iterations (msg123_set_control)")
    ap.add_argument("--no-mcilu", action="store_true",
                    help="skip the FP32 multicolour ILU(0) path")
    ap.add_argument("--no-fp64-dilu", action="store_true",
                    help="skip the mixed path's Newton options over the FP64 D-ILU")
    ap.add_argument("--kernel-groups", action="store_true",
                    help="time every kernel group with HIP events (slows the timed runs)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain 512 MiB device copy (HBM counter calibration)")
    args = ap.parse_args()
    global KERNEL_GROUPS
    KERNEL_GROUPS = args.kernel_groups
    if args.gpus != 1:
        print(json.dumps({"error": "the solver runs one domain on one GPU; --gpus must be 1"}))
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
    if args.config == "config3":
        print(json.dumps(ensemble_bench(args)), flush=True)
        return
    if args.config == "config2":
        print(json.dumps(seasonal_bench(args)), flush=True)
        return
    nsteps = args.warmup + args.steps
    stress = args.config == "config4"
    if stress:
        # the 134M-cell stress case: plain and mixed paths only (the network
        # was trained at dz = 0.25 m; the AMG's host aggregation of 134M cells
        # takes minutes), device memory reported
        import torch
        free0, total = torch.cuda.mem_get_info()
        print(f"config4: generating inputs", file=sys.stderr, flush=True)
    p = getattr(configs, args.config)(nsteps=nsteps) if args.config in ("config1", "config4") \
        else getattr(configs, args.config)()
    box = gpu.BoxMesh(p)
    L = box.b.mesh.len_scal
    if stress:
        print(f"config4: mesh built", file=sys.stderr, flush=True)

    # The reference build's own run of config 1 (tests/golden/ref_c1.npz: its
    # 48 accepted step times, Newton counts and the heads of 4095 sample
    # cells after every step): every path replays the reference's step times
    # and its heads are compared with the reference's at those cells.
    ref = None
    fpath = os.path.join(ROOT, "tests", "golden", "ref_c1.npz")
    if args.config == "config1" and os.path.exists(fpath):
        ref = dict(np.load(fpath))
        if nsteps > len(ref["step_times"]):
            ref = None
    ref_sched = None if ref is None else ref["step_times"][:nsteps].astype(np.float32)
    probes = None if ref is None else ref["sample_idx"]

    # plain FP64 (the reference's algorithm)
    plain = gpu.Context(box, precon=gpu.PRECON_DILU)
    n, nconn = plain.n, plain.nconn
    if stress:
        used_ctx = free0 - torch.cuda.mem_get_info()[0]
        print(f"config4: context created ({used_ctx / 1e9:.1f} GB)", file=sys.stderr, flush=True)
    res_p, wall_p, kern_p, warm_p = timed_run(plain, args.warmup, args.steps, schedule=ref_sched,
                                              probes=probes)
    sim_s = float(res_p["dts"].sum())
    newton_p = int(res_p["newton"].sum())
    s_plain = summary(res_p, wall_p, n, plain)
    s_plain["launch_shapes"] = plain.tune
    if ref is not None:
        s_plain["heads_vs_reference"] = vs_reference(plain, res_p, ref, args.warmup)
        nr = ref["step_newton"][args.warmup:nsteps]
        s_plain["newton_vs_reference"] = {"this_path": res_p["newton"].tolist(),
                                          "reference": nr.tolist()}
    value_p = s_plain["sim_days_per_hour"]
    _say(f"{args.config} plain: {value_p:.2f} sim-days/hour, {wall_p:.2f} s")
    spmv_ms, spmv_n = kern_p["spmv"]
    spmv_avg_ms = spmv_ms / max(spmv_n, 1)
    achieved = spmv_bytes(n, nconn) / (spmv_avg_ms * 1e-3) / 1e9
    h_plain = plain.get_state()["head_new"]
    # the step times the plain run took (warm-up + timed; the reference's
    # when the fixture is there), replayed by the other paths
    sched = np.concatenate([warm_p, res_p["times"]]).astype(np.float32)
    plain.close()

    traffic = None
    # the counter run measured config 1's SpMV: its figure is not this grid's
    tpath = os.path.join(ROOT, "profiles", "spmv_traffic_c1.json")
    if not stress and os.path.exists(tpath):
        traffic = json.load(open(tpath)).get("bytes_per_launch")
    out = {
        "metric": "sim-days/hour + SpMV GB/s at 4.2M cells: mixed+learned vs plain FP64, 1 GPU",
        "value": value_p, "unit": "sim-days/hour", "n_gpus": 1, "steps": int(res_p["nstep"]),
        "warmup": args.warmup, "ms_per_step": s_plain["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded GRF soil, config 1 statistics)",
        "config": {"workload": f"{args.config}: {p.nx}x{p.ny}x{p.nz} cells, {args.warmup}+"
                   f"{args.steps} accepted time steps on "
                   + ("the reference build's own step times" if ref is not None
                      else "the plain path's step times"),
                   "cells": n, "sim_seconds": sim_s,
                   "bottom": "free drainage" if p.free_drainage else "closed", "plain": s_plain},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_spmv_dot (SpMV fused with one dot, FP64, plain run)",
                     "bytes_per_launch": spmv_bytes(n, nconn), "avg_ms": spmv_avg_ms,
                     "launches": spmv_n,
                     # what the fused kernel itself must move: all seven diagonals
                     # (face zeros share cache lines), z and r-hat read, v written
                     "fused_kernel_bytes": 80 * n,
                     "fused_kernel_frac": 80 * n / (spmv_avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "kernel_ms": {"plain": {g: round(v[0], 3) for g, v in kern_p.items() if v[1]}},
    }

    def mixed_path(name, precon, warm=gpu.WARM_PERSIST, net=None, newton=MIXED_NEWTON):
        ctx = gpu.Context(box, precon=precon, consistent_jacobian=True)
        ctx.set_newton(**newton)
        ctx.set_law(gpu.LAW_LOGEXP)
        ctx.set_control({"host": gpu.CONTROL_HOST, "device": gpu.CONTROL_DEVICE,
                         "graph": gpu.CONTROL_GRAPH}[args.control])
        if net is not None:
            ctx.set_warm_net(net, np.log(p.ks), p.recharge_mmh * 1e-3 / 3600.0)
        # the SpMV roofline is the plain run's: no events in these timed runs
        res, wall, kern, _ = timed_run(ctx, args.warmup, args.steps, schedule=sched,
                                       warm_start=warm, probes=probes, spmv_events=False)
        h = ctx.get_state()["head_new"]
        s = summary(res, wall, n, ctx)
        s["same_step_times_as_plain"] = bool(np.array_equal(res["times"], res_p["times"]))
        s["max_abs_head_diff_vs_plain_m"] = float(np.abs(h - h_plain).max() * L)
        s["cells_beyond_1e-6_m_of_plain"] = int(np.count_nonzero(np.abs(h - h_plain) * L > 1e-6))
        s["speedup_vs_plain"] = s["sim_days_per_hour"] / value_p
        s["launch_shapes"] = ctx.tune
        if ref is not None:
            s["heads_vs_reference"] = vs_reference(ctx, res, ref, args.warmup)
        if warm == gpu.WARM_LEARNED:
            s["warm_start"] = ctx.warm_stats()
            # trial steps solved from the network's start and then again from
            # persistence: their Newton iterations are inside the wall time
            s["newton_iterations_incl_trials"] = (s["newton_iterations"]
                                                  + s["warm_start"]["shadow_newton"])
        if precon == gpu.PRECON_AMG32:
            st, fb = ctx.amg_stats()
            s.update(amg_setups=st, solves_falling_back_to_line32=fb,
                     warmup_wall_s_incl_hierarchy_setup=ctx.warmup_wall)
        ctx.close()
        out["config"][name] = s
        if KERNEL_GROUPS:
            out["kernel_ms"][name] = {g: round(v[0], 3) for g, v in kern.items() if v[1]}
        _say(f"{args.config} {name}: {s['sim_days_per_hour']:.2f} sim-days/hour, {wall:.2f} s")
        return s, h

    if stress:
        out["config"]["device_memory_per_context_gb"] = used_ctx / 1e9
        out["config"]["device_memory_total_gb"] = total / 1e9
        s_mixed, _ = mixed_path("mixed", gpu.PRECON_LINE32)
        out["value"] = s_mixed["sim_days_per_hour"]
        out["ms_per_step"] = s_mixed["ms_per_step"]
        out["dtype"] = "f64 (FP32 preconditioner)"
        out["metric"] = "sim-days/hour at 134M cells (config 4 stress case): mixed vs plain FP64, 1 GPU"
        args.no_cpu_baseline = True
    elif not args.plain_only:
        s_mixed, h_mixed = mixed_path("mixed", gpu.PRECON_LINE32)
        net = gpu.Net(p.nx, p.ny, p.nz)
        net.params(np.load(WEIGHTS))
        s_learned, h_learned = mixed_path("mixed_learned", gpu.PRECON_LINE32, gpu.WARM_LEARNED,
                                          net)
        net.close()
        s_learned["gain_vs_mixed_without_network"] = (s_learned["sim_days_per_hour"]
                                                      / s_mixed["sim_days_per_hour"])
        # the value path's most frequent stencil kernel, F(h) (about 1.6 calls
        # per Newton iteration): its average duration from HIP events over a
        # short untimed run of the mixed path with every kernel group timed
        rctx = gpu.Context(box, precon=gpu.PRECON_LINE32, consistent_jacobian=True)
        rctx.set_newton(**MIXED_NEWTON)
        rctx.set_law(gpu.LAW_LOGEXP)
        rctx.run(nstep_max=2, keep_heads=False, schedule=sched)
        rctx.autotune(3)
        rctx.set_timing(True)
        rctx.run(nstep_max=max(min(8, nsteps) - 2, 1), keep_heads=False, reset=False,
                 schedule=sched)
        l_ms, l_cnt = rctx.timing("line_apply")
        a_ms, a_cnt = rctx.timing("assemble")
        # F(h) alone: the Newton iteration now evaluates it only when the line
        # search backtracks (the next system, assembled ahead, gives F at the
        # full step), so the residual kernel is timed on the run's last head
        rctx.set_timing(False)
        rctx.set_timing(True)
        h_last = rctx.get_state()["head_new"]
        for _ in range(10):
            rctx.calc_func(h_last)
        r_ms, r_cnt = rctx.timing("residual")
        rctx.close()
        r_avg = r_ms / max(r_cnt, 1)
        r_bytes = residual_bytes(n, p.nx * p.ny)
        r_traffic = None
        rpath = os.path.join(ROOT, "profiles", "residual_traffic_c1.json")
        if os.path.exists(rpath):
            r_traffic = json.load(open(rpath)).get("bytes_per_launch")
        r_ach = r_bytes / (r_avg * 1e-3) / 1e9 if r_cnt else 0.0
        out["roofline_residual"] = {
            "bound": "hbm", "achieved": r_ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": r_ach / HBM_PEAK_GBS, "traffic": r_traffic,
            "kernel": "k_fused<kResid> (F(h) with ||F||^2, law on chip, mixed path)",
            "bytes_per_launch": r_bytes, "avg_ms": r_avg, "launches": r_cnt}
        a_avg = a_ms / max(a_cnt, 1)
        a_bytes = assembly_bytes(n, p.nx * p.ny, s_a=4)
        a_ach = a_bytes / (a_avg * 1e-3) / 1e9 if a_cnt else 0.0
        a_traffic = None
        apath = os.path.join(ROOT, "profiles", "assembly_traffic_c1.json")
        if os.path.exists(apath):
            a_traffic = json.load(open(apath)).get("bytes_per_launch")
        out["roofline_assembly"] = {
            "bound": "hbm", "achieved": a_ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": a_ach / HBM_PEAK_GBS, "traffic": a_traffic,
            "kernel": "k_fused<kAssem> (rhs and FD Jacobian in FP32 diagonals, law on chip, "
                      "mixed path)",
            "bytes_per_launch": a_bytes, "avg_ms": a_avg, "launches": a_cnt}
        # the mixed path's smoother: one FP32 tridiagonal solve per column
        # (k_zline_apply), r read in FP64, three FP32 factor arrays, z written
        l_avg = l_ms / max(l_cnt, 1)
        l_bytes = 28 * n
        l_ach = l_bytes / (l_avg * 1e-3) / 1e9 if l_cnt else 0.0
        l_traffic = None
        lpath = os.path.join(ROOT, "profiles", "line_precond_traffic_c1.json")
        if os.path.exists(lpath):
            l_traffic = json.load(open(lpath)).get("bytes_per_launch")
        out["roofline_line_precond"] = {
            "bound": "hbm", "achieved": l_ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": l_ach / HBM_PEAK_GBS, "traffic": l_traffic,
            "kernel": "k_zline_apply (FP32 vertical-line preconditioner, mixed path)",
            "bytes_per_launch": l_bytes, "avg_ms": l_avg, "launches": l_cnt}
        out["value"] = s_learned["sim_days_per_hour"]
        out["ms_per_step"] = s_learned["ms_per_step"]
        ws = s_learned["warm_start"]
        if ws["net_on"] == 1 and ws["attempts"] > 0:
            out["dtype"] = "f64 (FP32 preconditioner, bf16 network)"
            out["value_path"] = "mixed + learned warm start (network in use)"
        else:
            # the gate declined the network: what ran after its trials is the
            # mixed path (the trial steps' cost is inside the timed region)
            out["dtype"] = "f64 (FP32 preconditioner)"
            out["value_path"] = (f"mixed path: the learned start's gate declined the network "
                                 f"after {ws['shadows']} trial steps "
                                 f"({ws['shadow_newton']} trial Newton iterations, timed)")
        if not args.no_fp64_dilu:
            # the mixed path's Newton options, law and Jacobian over the plain
            # path's FP64 D-ILU: what the FP32 line preconditioner adds
            s64, _ = mixed_path("mixed_fp64_dilu", gpu.PRECON_DILU)
            s64["speedup_of_fp32_line_over_this"] = (s_mixed["sim_days_per_hour"]
                                                     / s64["sim_days_per_hour"])
        if not args.no_mcilu:
            mixed_path("mixed_mcilu", gpu.PRECON_MCILU32)
            # its apply (two colour passes) timed with HIP events over a short
            # untimed run: 7 FP32 factors per cell, r read and z written in
            # FP64, the black half of z read again by the red pass
            mctx = gpu.Context(box, precon=gpu.PRECON_MCILU32, consistent_jacobian=True)
            mctx.set_newton(**MIXED_NEWTON)
            mctx.set_law(gpu.LAW_LOGEXP)
            mctx.run(nstep_max=2, keep_heads=False, schedule=sched)
            mctx.set_timing(True)
            mctx.run(nstep_max=max(min(6, nsteps) - 2, 1), keep_heads=False, reset=False,
                     schedule=sched)
            m_ms, m_cnt = mctx.timing("mc_apply")
            f_ms, f_cnt = mctx.timing("mc_factor")
            mctx.close()
            m_avg = m_ms / max(m_cnt, 1)
            m_bytes = 48 * n
            m_ach = m_bytes / (m_avg * 1e-3) / 1e9 if m_cnt else 0.0
            m_traffic = None
            mpath = os.path.join(ROOT, "profiles", "mcilu_traffic_c1.json")
            if os.path.exists(mpath):
                mj = json.load(open(mpath))
                m_traffic = mj["black"]["bytes_per_launch"] + mj["red"]["bytes_per_launch"]
            out["roofline_mcilu_apply"] = {
                "bound": "hbm", "achieved": m_ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": m_ach / HBM_PEAK_GBS, "traffic": m_traffic,
                "kernel": "k_mc_black + k_mc_red (FP32 multicolour ILU(0) apply)",
                "bytes_per_launch": m_bytes, "avg_ms": m_avg, "launches": m_cnt,
                "factor_avg_ms": f_ms / max(f_cnt, 1)}
        if not args.no_amg:
            mixed_path("mixed_amg", gpu.PRECON_AMG32, newton=AMG_NEWTON)

    if not args.no_cpu_baseline:
        # the first accepted step: the reference's (dt 18.75 s) or the plain run's
        dt0 = float(ref["step_dts"][0]) if ref is not None else float(
            np.concatenate([np.diff(np.concatenate([[0.0], warm_p])), res_p["dts"]])[0])
        out["cpu_baseline"] = cpu_baseline(p, dt0)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
