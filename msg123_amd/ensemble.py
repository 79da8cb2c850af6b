"""Many independent domains at once: one context (and HIP stream) per
domain, driven by a pool of host threads so that the GPU overlaps the
domains' kernels.  Every domain's arithmetic is exactly its standalone
solve, so the results are bit-identical to solving the domains one after
another.

    from msg123_amd import configs, ensemble
    res = ensemble.run(configs.config3(), threads=16)

Needs a GPU.
"""
from __future__ import annotations

import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from msg123_amd import gpu

# the mixed path's options (as bench.py runs them)
MIXED = dict(precon=gpu.PRECON_LINE32, consistent_jacobian=True)


def _solve(p, kw, newton, law, keep_heads, nstep_max):
    t0 = time.perf_counter()
    box = gpu.BoxMesh(p)
    ctx = gpu.Context(box, **kw)
    if newton:
        ctx.set_newton(**newton)
    if law is not None:
        ctx.set_law(law)
    t1 = time.perf_counter()
    res = ctx.run(nstep_max=nstep_max, keep_heads=keep_heads)
    t2 = time.perf_counter()
    ctx.close()
    return dict(name=p.name, cells=p.ncell, steps=int(res["nstep"]),
                newton=int(res["newton"].sum()), setup_s=t1 - t0, solve_s=t2 - t1,
                heads=res["heads"] if keep_heads else None, times=res["times"])


def run(problems, threads=16, mixed=True, keep_heads=False, nstep_max=100000):
    """Solve every problem with the reference's time control; returns
    per-instance results (solve wall time measured inside its thread) and
    the ensemble's totals."""
    kw = dict(MIXED) if mixed else dict(precon=gpu.PRECON_DILU)
    newton = gpu.MIXED_NEWTON if mixed else None
    law = gpu.LAW_LOGEXP if mixed else None
    t0 = time.perf_counter()
    if threads <= 1:
        out = [_solve(p, kw, newton, law, keep_heads, nstep_max) for p in problems]
    else:
        # largest first (longest-processing-time order), so that the big
        # instances do not start last and leave the pool idle behind them
        order = sorted(range(len(problems)), key=lambda k: -problems[k].ncell)
        with ThreadPoolExecutor(max_workers=threads) as ex:
            done = list(ex.map(lambda k: _solve(problems[k], kw, newton, law, keep_heads,
                                                nstep_max), order))
        out = [None] * len(problems)
        for k, r in zip(order, done):
            out[k] = r
    wall = time.perf_counter() - t0
    solve = np.array([r["solve_s"] for r in out])
    steps = sum(r["steps"] for r in out)
    return dict(instances=out, wall_s=wall, instance_steps=steps,
                instance_steps_per_s=steps / wall,
                cells=sum(r["cells"] for r in out),
                solve_p50_s=float(np.percentile(solve, 50)),
                solve_p99_s=float(np.percentile(solve, 99)))
