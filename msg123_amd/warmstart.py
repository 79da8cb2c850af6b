"""Training of the learned Newton warm start on solver trajectories.

The solver itself generates the data: for each training seed it runs the
time loop on a box with config-1 statistics (GRF lnKs, four layers,
recharge from the top) and keeps every accepted head.  Each step k >= 1 gives
one sample: features from (h_{k-1}, h_k, dt ratio, recharge depth, lnKs) and
the target (h_{k+1} - h_lin) / scale, all built on the device by the
library's feature kernels.  Training is Adam on the device kernels
(msg123_net_train_step); the result is the fp32 master parameter vector.

    python -m msg123_amd.warmstart --out msg123_amd/warmstart_c1.npy

Needs a GPU.  Device buffers are torch tensors (memory only; every
computation runs in libmsg123_gpu.so).
"""
from __future__ import annotations

import argparse
import time

import numpy as np

from msg123_amd import configs, gpu


def trajectory(p, nsteps, precon=gpu.PRECON_LINE32):
    """Accepted heads (scaled) of one run, with the initial head first, and the
    step lengths."""
    box = gpu.BoxMesh(p)
    ctx = gpu.Context(box, precon=precon, consistent_jacobian=(precon == gpu.PRECON_LINE32))
    res = ctx.run(nstep_max=nsteps)
    ctx.close()
    h0 = box.arrays()["h0"]
    return box, np.vstack([h0[None, :], res["heads"]]), res["dts"]


class Samples:
    """Training samples of one trajectory, resident in HBM."""

    def __init__(self, net, box, heads, dts, rech_ms):
        import torch
        a = box.arrays()
        n = a["n"]
        L = a["len_scal"]
        dev = "cuda"
        top = torch.from_numpy(a["cell_top"]).to(dev)
        lnks = torch.from_numpy(np.log(box.p.ks).astype(np.float32)).to(dev)
        hs = [torch.from_numpy(h).to(dev) for h in heads]
        scale = torch.zeros(1, dtype=torch.float64, device=dev)
        hlin = torch.zeros(n, dtype=torch.float64, device=dev)
        self.x, self.t = [], []
        # heads[k] is the state after k accepted steps; dts[k] is the length of step k+1
        for k in range(1, len(heads) - 1):
            r = dts[k] / dts[k - 1]
            x = torch.zeros(n * 4, dtype=torch.int16, device=dev)
            t = torch.zeros(n, dtype=torch.float32, device=dev)
            net.features(hs[k], hs[k - 1], top, lnks, r, L, rech_ms * dts[k], x, hlin, scale)
            net.target(hs[k + 1], hlin, scale, L, t)
            self.x.append(x)
            self.t.append(t)


def train(net, samples, epochs, lr0, seed=0, log=print):
    """Adam over all samples for `epochs` passes, cosine-decayed learning rate."""
    pairs = [(x, t) for s in samples for x, t in zip(s.x, s.t)]
    rng = np.random.default_rng(seed)
    total = epochs * len(pairs)
    k = 0
    for ep in range(epochs):
        order = rng.permutation(len(pairs))
        losses = []
        for i in order:
            lr = lr0 * 0.5 * (1.0 + np.cos(np.pi * k / total))
            losses.append(net.train_step(pairs[i][0], pairs[i][1], max(lr, 1e-6)))
            k += 1
        log(f"epoch {ep + 1}/{epochs}: mean loss {np.mean(losses):.4e}")
    return float(np.mean(losses))


def evaluate(p, params, nsteps, log=print):
    """Newton iterations and wall time of the mixed solver on `p` with the
    persistence, linear and learned warm starts, all replaying the step
    times of the persistence run; heads checked against it."""
    box = gpu.BoxMesh(p)
    kw = dict(precon=gpu.PRECON_LINE32, consistent_jacobian=True)
    out = {}
    base = None
    net = None
    for mode in (gpu.WARM_PERSIST, gpu.WARM_LINEAR, gpu.WARM_LEARNED):
        ctx = gpu.Context(box, **kw)
        if mode == gpu.WARM_LEARNED:
            net = gpu.Net(p.nx, p.ny, p.nz)
            net.params(params)
            ctx.set_warm_net(net, np.log(p.ks), p.recharge_mmh * 1e-3 / 3600.0)
        sched = None if base is None else base["times"].astype(np.float32)
        t0 = time.perf_counter()
        res = ctx.run(nstep_max=nsteps, schedule=sched, warm_start=mode)
        wall = time.perf_counter() - t0
        if base is None:
            base = res
        same = bool(np.array_equal(res["times"], base["times"]))
        err = float(np.abs(res["heads"][-1] - base["heads"][-1]).max() * box.b.mesh.len_scal)
        name = ("persist", "linear", "learned")[mode]
        out[name] = dict(newton=int(res["newton"].sum()), wall_s=wall, steps=int(res["nstep"]),
                         same_step_times=same, max_head_diff_m=err, warm=ctx.warm_stats(),
                         newton_per_step=res["newton"].tolist())
        log(f"{name}: {out[name]}")
        ctx.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default="msg123_amd/warmstart_c1.npy")
    ap.add_argument("--seeds", default="100,101,102,103,104,105,106,107")
    ap.add_argument("--grid", default="128,128,64")
    ap.add_argument("--steps", type=int, default=48)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--eval", default="", help="config-1 seeds to evaluate on afterwards")
    args = ap.parse_args()
    # torch (device buffers) must bring up the GPU before the library does
    import torch
    torch.zeros(1, device="cuda")
    nx, ny, nz = (int(v) for v in args.grid.split(","))
    seeds = [int(s) for s in args.seeds.split(",")]
    net = gpu.Net(nx, ny, nz, seed=2024)
    samples = []
    t0 = time.perf_counter()
    for s in seeds:
        p = configs.make_grf(f"train{s}", nx, ny, nz, 5.0, 5.0, 0.25, s, nsteps=args.steps)
        box, heads, dts = trajectory(p, args.steps)
        samples.append(Samples(net, box, heads, dts, p.recharge_mmh * 1e-3 / 3600.0))
        print(f"seed {s}: {len(dts)} steps, {len(samples[-1].x)} samples "
              f"({time.perf_counter() - t0:.1f} s)", flush=True)
    t1 = time.perf_counter()
    loss = train(net, samples, args.epochs, args.lr, log=lambda m: print(m, flush=True))
    print(f"trained in {time.perf_counter() - t1:.1f} s, final loss {loss:.4e}", flush=True)
    np.save(args.out, net.params())
    if args.eval:
        params = net.params()
        net.close()
        for s in (int(v) for v in args.eval.split(",")):
            evaluate(configs.config1(seed=s, nsteps=args.steps), params, args.steps,
                     log=lambda m: print(f"eval seed {s}: {m}", flush=True))


if __name__ == "__main__":
    main()
