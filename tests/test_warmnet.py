"""The warm-start network: the CPU oracle (oracle/warmnet.py) pinned by loops
and finite differences; the HIP forward/backward/Adam kernels against it.
Tolerances of the GPU tests come from bf16 storage: the oracle rounds at the
same points, so the two differ by fp32-vs-fp64 accumulation and the bf16
roundings that difference flips."""
import numpy as np
import pytest

from oracle import warmnet as W


def _rand_params(seed, scale_last=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(W.NPARAM)
    for l, (wo, bo) in enumerate(W.LAYOUT):
        lim = np.sqrt(6.0 / (27 * W.CIN[l])) * (scale_last if l == 3 else 1.0)
        p[wo:bo] = rng.uniform(-lim, lim, bo - wo)
        p[bo:bo + W.COUT[l]] = rng.uniform(-0.1, 0.1, W.COUT[l])
    return p.astype(np.float32).astype(np.float64)


def _rand_x(shape, seed):
    rng = np.random.default_rng(seed)
    return W.bf16(rng.normal(size=shape + (4,)))


def test_oracle_conv_matches_loops():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(3, 4, 5, 2))
    w = rng.normal(size=(27, 3, 2))
    y = W.conv(x, w)
    nz, ny, nx = x.shape[:3]
    ref = np.zeros((nz, ny, nx, 3))
    for z in range(nz):
        for yy in range(ny):
            for xx in range(nx):
                for t, (dz, dy, dx) in enumerate(W.TAPS):
                    zz, y2, x2 = z + dz, yy + dy, xx + dx
                    if 0 <= zz < nz and 0 <= y2 < ny and 0 <= x2 < nx:
                        ref[z, yy, xx] += w[t] @ x[zz, y2, x2]
    np.testing.assert_allclose(y, ref, rtol=1e-12, atol=1e-12)


def test_oracle_gradient_matches_finite_differences():
    """FP64 gradient (no bf16 rounding) against central differences."""
    p = _rand_params(1, scale_last=1.0)
    rng = np.random.default_rng(2)
    x = rng.normal(size=(3, 3, 4, 4))
    t = rng.normal(size=(3, 3, 4))
    _, g = W.loss_grad(p, x, t, emulate=False)
    idx = []
    for l, (wo, bo) in enumerate(W.LAYOUT):
        idx += list(rng.choice(np.arange(wo, bo), 6, replace=False)) + [bo]
    for i in idx:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (W.loss_grad(pp, x, t, False)[0] - W.loss_grad(pm, x, t, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -9, -2.5e-3, 3.0e5])
    r = W.bf16(vals)
    # bf16 spacing at 1 is 2^-7: 1 + 2^-8 is the halfway point and ties go to
    # the even mantissa (1), 1 + 2^-9 rounds down, 1 + 3*2^-9 rounds up
    assert r[0] == 1.0 and r[1] == 1.0 and r[3] == 1.0
    assert r[2] == 1.0 + 2 ** -7
    assert np.all(np.abs(r - vals) <= np.abs(vals) * 2 ** -8)


# ---------------------------------------------------------------------------
# GPU


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _dev_bits(torch, x):
    bits = W.bf16_bits(x.reshape(-1, 4)).view(np.int16)
    return torch.from_numpy(bits.copy()).cuda()


@pytest.mark.gpu
def test_net_initial_parameters_match_oracle():
    from msg123_amd import gpu
    _torch()
    net = gpu.Net(5, 4, 3, seed=77)
    np.testing.assert_array_equal(net.params().astype(np.float64), W.init_params(77))


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(5, 6, 7), (4, 8, 16), (3, 5, 33), (6, 12, 48), (5, 8, 32)])
def test_net_forward_vs_oracle(shape):
    """MFMA forward (bf16, fp32 accumulate) against the oracle on grids whose
    16-cell tiles cross rows and planes, and (nx a multiple of 16, ny of 4)
    on the LDS-staged 16 x 4 tiles, several per row and column, two levels
    per block (nz even) or one (nz odd)."""
    torch = _torch()
    from msg123_amd import gpu
    nz, ny, nx = shape
    p = _rand_params(3, scale_last=3.0)
    x = _rand_x(shape, 4)
    net = gpu.Net(nx, ny, nz)
    net.params(p)
    xd = _dev_bits(torch, x)
    yd = torch.zeros(nx * ny * nz, dtype=torch.float32, device="cuda")
    net.forward(xd, yd)
    y_ref, _ = W.forward(p, x, emulate=True)
    y = yd.cpu().numpy().reshape(shape)
    err = np.abs(y - y_ref).max() / np.abs(y_ref).max()
    print("forward rel err", err, flush=True)
    assert err <= 2e-2


@pytest.mark.gpu
def test_net_gradient_vs_oracle():
    """Backward (input gradients on the mirrored-tap kernel, weight gradients
    on MFMA over the cells) against the oracle, per layer."""
    torch = _torch()
    from msg123_amd import gpu
    shape = (4, 6, 10)
    nz, ny, nx = shape
    p = _rand_params(5, scale_last=3.0)
    x = _rand_x(shape, 6)
    t = np.random.default_rng(7).normal(size=shape)
    net = gpu.Net(nx, ny, nz)
    net.params(p)
    loss, g = net.grad(_dev_bits(torch, x), torch.from_numpy(t.astype(np.float32).ravel()).cuda())
    loss_ref, g_ref = W.loss_grad(p, x, t.astype(np.float32).astype(np.float64), emulate=True)
    assert abs(loss - loss_ref) <= 1e-3 * loss_ref
    for l, (wo, bo) in enumerate(W.LAYOUT):
        seg = slice(wo, bo + W.COUT[l])
        rel = np.linalg.norm(g[seg] - g_ref[seg]) / np.linalg.norm(g_ref[seg])
        print("layer", l, "grad rel err", rel, flush=True)
        assert rel <= 3e-2, (l, rel)


@pytest.mark.gpu
def test_net_features_vs_oracle():
    torch = _torch()
    from msg123_amd import gpu
    nx, ny, nz = 9, 7, 5
    n = nx * ny * nz
    rng = np.random.default_rng(8)
    hn = rng.normal(size=n) * 0.1 - 1.0
    hm = hn - rng.normal(size=n) * 1e-3
    top = rng.normal(size=n) * 0.01
    lnks = np.log(10 ** rng.uniform(-7, -3, n)).astype(np.float32)
    net = gpu.Net(nx, ny, nz)
    dd = {k: torch.from_numpy(v).cuda() for k, v in dict(hn=hn, hm=hm, top=top, lnks=lnks).items()}
    x = torch.zeros(n * 4, dtype=torch.int16, device="cuda")
    hlin = torch.zeros(n, dtype=torch.float64, device="cuda")
    sc = torch.zeros(1, dtype=torch.float64, device="cuda")
    net.features(dd["hn"], dd["hm"], dd["top"], dd["lnks"], 1.2, 10.0, 7.2e-4, x, hlin, sc)
    f_ref, hlin_ref, s_ref = W.features(hn, hm, top, lnks, 1.2, 10.0, 7.2e-4, nx * ny)
    f = W.from_bits(x.cpu().numpy().view(np.uint16)).reshape(n, 4)
    assert abs(sc.item() - s_ref) <= 1e-12 * s_ref
    np.testing.assert_array_equal(hlin.cpu().numpy(), hlin_ref)
    np.testing.assert_allclose(f, f_ref, rtol=2 ** -7, atol=0)


@pytest.mark.gpu
def test_net_training_fits_a_teacher():
    """Adam on the device kernels drives the loss against a fixed teacher
    network's outputs down by an order of magnitude."""
    torch = _torch()
    from msg123_amd import gpu
    shape = (8, 16, 16)
    nz, ny, nx = shape
    teacher = _rand_params(11, scale_last=1.0)
    xs = [_rand_x(shape, 100 + k) for k in range(4)]
    ts = [W.forward(teacher, x, emulate=True)[0] for x in xs]
    net = gpu.Net(nx, ny, nz, seed=5)
    xd = [_dev_bits(torch, x) for x in xs]
    td = [torch.from_numpy(t.astype(np.float32).ravel()).cuda() for t in ts]
    first = np.mean([net.grad(a, b)[0] for a, b in zip(xd, td)])
    for it in range(300):
        net.train_step(xd[it % 4], td[it % 4], 3e-3)
    last = np.mean([net.grad(a, b)[0] for a, b in zip(xd, td)])
    print("teacher loss", first, "->", last, flush=True)
    assert last <= 0.1 * first


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["linear", "select", "learned"])
def test_warm_start_keeps_the_answer(mode):
    """A different Newton initial head changes iteration counts, not the
    converged heads: replaying the persistence run's step times, every step
    lands within 1e-6 m of it.  (The learned mode here uses an untrained
    network; its gate tries it on one step at least, cuts a trial off at the
    Newton count of the last step accepted from persistence and then
    declines the network.)"""
    torch = _torch()
    from msg123_amd import configs, gpu
    p = configs.small(12, 10, 8, seed=5, nsteps=6)
    box = gpu.BoxMesh(p)
    base = gpu.Context(box).run()
    ctx = gpu.Context(box)
    if mode == "learned":
        net = gpu.Net(p.nx, p.ny, p.nz, seed=1)
        ctx.set_warm_net(net, np.log(p.ks), p.recharge_mmh * 1e-3 / 3600.0)
    warm = {"linear": gpu.WARM_LINEAR, "select": gpu.WARM_SELECT, "learned": gpu.WARM_LEARNED}
    res = ctx.run(schedule=base["times"].astype(np.float32), warm_start=warm[mode])
    np.testing.assert_array_equal(res["times"], base["times"])
    err = np.abs(res["heads"] - base["heads"]).max() * box.b.mesh.len_scal
    print(mode, "newton", res["newton"].tolist(), "vs", base["newton"].tolist(), "err", err,
          ctx.warm_stats(), flush=True)
    assert err <= 1e-6
    if mode == "learned":
        st = ctx.warm_stats()
        assert 1 <= st["attempts"] <= len(res["times"]) - 1
        assert st["shadows"] <= 2
        if st["net_on"] == 0 and st["shadows"] == 1 and st["attempts"] == 1:
            # the one trial was cut off: at most one accepted step's count
            assert st["shadow_newton"] <= int(res["newton"].max())
        if st["shadows"] == st["attempts"]:
            # every network start was a trial, discarded: persistence's run
            np.testing.assert_array_equal(res["heads"], base["heads"])
    del torch


@pytest.mark.gpu
def test_trained_network_gain():
    """What the trained network (msg123_amd/warmstart_c1.npy, trained on
    config-1-statistics trajectories of seeds 100-107 at 128x128x64) gains
    over the same solver without it, on a held-out seed and grid (64x64x32,
    seed 9): Newton iterations of the mixed path with the persistence start
    vs the learned start, both replaying the step times of the plain FP64
    path (the bench's protocol).  The answer is the same; the gate bounds what
    a network that does not help costs (two trial steps of at most 20 Newton
    iterations each, solved again from persistence)."""
    torch = _torch()
    import os
    from msg123_amd import configs, gpu
    nsteps = 24
    p = configs.make_grf("heldout9", 64, 64, 32, 5.0, 5.0, 0.25, 9, nsteps=nsteps)
    box = gpu.BoxMesh(p)
    plain = gpu.Context(box, precon=gpu.PRECON_DILU)
    sched = plain.run(nstep_max=nsteps)["times"].astype(np.float32)
    plain.close()

    def mixed(net=None):
        ctx = gpu.Context(box, precon=gpu.PRECON_LINE32, consistent_jacobian=True)
        ctx.set_newton(0.1, 0.1, True, False)
        if net is not None:
            ctx.set_warm_net(net, np.log(p.ks), p.recharge_mmh * 1e-3 / 3600.0)
        r = ctx.run(nstep_max=nsteps, schedule=sched,
                    warm_start=gpu.WARM_PERSIST if net is None else gpu.WARM_LEARNED)
        st = ctx.warm_stats()
        ctx.close()
        return r, st

    base, _ = mixed()
    net = gpu.Net(p.nx, p.ny, p.nz)
    net.params(np.load(os.path.join(os.path.dirname(os.path.dirname(__file__)), "msg123_amd",
                                    "warmstart_c1.npy")))
    res, st = mixed(net)
    net.close()
    np.testing.assert_array_equal(res["times"], base["times"])
    err = np.abs(res["heads"] - base["heads"]).max() * box.b.mesh.len_scal
    # the learned run's cost includes the discarded trial solves
    nb, nl = int(base["newton"].sum()), int(res["newton"].sum()) + st["shadow_newton"]
    print(f"held-out 64x64x32 seed 9, {nsteps} steps: Newton persistence {nb}, learned {nl} "
          f"incl. trials (gain {nb / nl:.3f}x), max|dh| {err:.2e} m, {st}", flush=True)
    assert err <= 1e-6
    assert st["shadows"] <= 2 and st["shadow_newton"] <= 2 * 20
    assert nl <= nb + 2 * 20
    del torch
