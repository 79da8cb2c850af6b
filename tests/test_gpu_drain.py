"""Free-drainage bottom (msg123_set_bottom_drain; no reference counterpart,
the reference's bottom is closed).  Pinned by closed forms, not by the
oracle: a uniform column with a uniform pressure head psi* and top recharge
q = Ks kr(psi*) is a steady state of the drained column (gravity flow
Ks kr A between every pair of layers, the same out of the bottom), so the
heads may not move; with the bottom closed the same column fills from the
bottom.  A transient column must close its mass balance with the outflow
in the connection term, to the Newton tolerance.  kr is the textbook van Genuchten-Mualem law,
evaluated here in float64."""
import sys

import numpy as np
import pytest

from msg123_amd import gpu
from msg123_amd.configs import Problem

pytestmark = pytest.mark.gpu

ALPHA, VGN, KS, PSI_STAR = 2.0, 1.6, 1.0e-5, -1.0


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _kr(psi):
    m = 1.0 - 1.0 / VGN
    se = (1.0 + (ALPHA * abs(psi)) ** VGN) ** (-m)
    return np.sqrt(se) * (1.0 - (1.0 - se ** (1.0 / m)) ** m) ** 2


def _column(psi0, free, nsteps=10, nz=30):
    nx = ny = 3
    n = nx * ny * nz
    q = KS * _kr(PSI_STAR)
    return Problem(
        name="drain", nx=nx, ny=ny, nz=nz, dx=1.0, dy=1.0, dz=0.1,
        alpha=np.full(n, ALPHA, np.float32), vg_n=np.full(n, VGN, np.float32),
        theta_r=np.full(n, 0.05, np.float32), theta_s=np.full(n, 0.40, np.float32),
        ks=np.full(n, KS, np.float32), ss=np.full(n, 1e-5, np.float32),
        wt_depth=10.0, recharge_mmh=q * 1e3 * 3600.0, init_step=3600.0, max_step=3600.0,
        end_time=3600.0 * nsteps, h0_pressure=psi0, free_drainage=free)


PATHS = {
    "plain": (dict(), None, None),
    "mixed": (dict(precon=gpu.PRECON_LINE32, consistent_jacobian=True), gpu.MIXED_NEWTON,
              gpu.LAW_LOGEXP),
}


def _run(p, path):
    kw, newton, law = PATHS[path]
    box = gpu.BoxMesh(p)
    ctx = gpu.Context(box, **kw)
    if newton:
        ctx.set_newton(**newton)
    if law is not None:
        ctx.set_law(law)
    ctx.set_mass_balance(True)
    res = ctx.run()
    mb = ctx.mass_balance()
    L = box.b.mesh.len_scal
    drift = np.abs(res["heads"][: res["nstep"]] - box.arrays()["h0"]).max() * L
    ctx.close()
    return res, mb, drift


@pytest.mark.parametrize("path", sorted(PATHS))
def test_drained_column_steady_state(path):
    _say("test_drained_column_steady_state", path)
    res, mb, drift = _run(_column(PSI_STAR, True), path)
    _say("  drained: max|h - h0| m", drift, "newton", res["newton"][: res["nstep"]].tolist())
    assert res["nstep"] == 10
    assert drift <= 1e-6
    # the outflow equals the recharge: q A per column over the 10 h
    assert mb["connection"] < 0
    assert abs(mb["connection"] + mb["recharge"]) <= 1e-5 * mb["recharge_abs"]
    res, mb, drift = _run(_column(PSI_STAR, False), path)
    _say("  closed: max|h - h0| m", drift)
    assert drift > 1e-3
    assert abs(mb["connection"]) <= 1e-10 * mb["recharge_abs"]


@pytest.mark.parametrize("path", sorted(PATHS))
def test_drained_column_mass_balance(path):
    _say("test_drained_column_mass_balance", path)
    res, mb, _ = _run(_column(-3.0, True, nsteps=24), path)
    _say("  ", {k: mb[k] for k in ("storage", "connection", "recharge", "residual", "total")})
    assert res["nstep"] >= 24
    assert mb["connection"] < 0  # water leaves through the bottom
    # the accounting's residual (with the outflow) is the Newton residual the
    # solve drove to its tolerance: without the outflow it would be the
    # drained volume, 3% of the recharge here.  1e-6 relative is the Newton
    # tolerance of so small a column (plain path 7e-9, mixed below 1e-10)
    assert abs(mb["residual"]) <= 1e-6 * mb["recharge_abs"]
    assert abs(mb["total"]) <= 1e-6 * mb["recharge_abs"]
    assert abs(mb["connection"]) >= 1e-2 * mb["recharge_abs"]


def test_bottom_drain_arguments():
    _say("test_bottom_drain_arguments")
    p = _column(PSI_STAR, False, nsteps=1, nz=4)
    box = gpu.BoxMesh(p)
    ctx = gpu.Context(box)
    try:
        cond = box.drain_conductance()
        with pytest.raises(ValueError):
            ctx.set_bottom_drain(cond[:-1])
        bad = cond.copy()
        bad[2] = -1.0
        with pytest.raises(RuntimeError):
            ctx.set_bottom_drain(bad)
        bad[2] = np.nan
        with pytest.raises(RuntimeError):
            ctx.set_bottom_drain(bad)
        ctx.set_bottom_drain(cond)
        ctx.set_bottom_drain(None)
    finally:
        ctx.close()


def test_config4_statistics_plain_and_mixed_agree():
    """Config 4's setting (dry soil at -10 m pressure head, 5e-6 m/s
    infiltration, GRF lnKs with sigma 2, n up to 3, free-drainage bottom) on
    a 128 x 128 x 32 box over 1800 s: the plain path (the reference's
    algorithm, which cuts the 600 s steps to 75 s at the sharp front) and the
    mixed path on the plain path's step times end within 1e-6 m of each other
    at every step, both conserve mass to 1e-9 of the water put in, and the
    wetting front has entered the top layer."""
    from msg123_amd import configs
    p = configs.config4(nsteps=3, dims=(128, 128, 32))
    box = gpu.BoxMesh(p)
    runs = []
    for mixed in (False, True):
        if mixed:
            ctx = gpu.Context(box, precon=gpu.PRECON_LINE32, consistent_jacobian=True)
            ctx.set_law(gpu.LAW_LOGEXP)
            ctx.set_newton(**gpu.MIXED_NEWTON)
        else:
            ctx = gpu.Context(box)
        ctx.set_mass_balance(True)
        res = ctx.run(schedule=runs[0]["times"] if mixed else None)
        res["mass"] = ctx.mass_balance()
        runs.append(res)
        ctx.close()
    a, b = runs
    L = box.b.mesh.len_scal
    print("steps", a["nstep"], b["nstep"], "newton", a["newton"].tolist(), b["newton"].tolist(),
          flush=True)
    np.testing.assert_array_equal(a["times"], b["times"])
    assert np.abs(a["heads"] - b["heads"]).max() * L <= 1e-6
    for r in runs:
        m = r["mass"]
        assert abs(m["residual"]) <= 1e-9 * m["recharge_abs"], m
    top = box.arrays()["cell_top"][:p.nx * p.ny]
    psi0 = -10.0 / L
    assert (a["heads"][-1][:p.nx * p.ny] - top).max() > psi0 + 1.0 / L  # wetter by > 1 m
