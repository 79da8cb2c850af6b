"""GPU parity of the river, lake, well and sea terms (msg123_set_bterms):
func_riveterm, func_laketerm, func_wellterm, func_sealterm
(Calc_function.f90:277-302,393-513), set_rivebound, set_lakebound,
set_seabound (Make_linearsystem.f90:287-408,499-547) and their calc_mass
terms, against the reference run that had them set in its module variables
(tests/golden/ref_bnd.npz, case "bnd": rivers and lakes on two rows of the
top layer over all three branches of their exchange, three wells, sea
connections on the x = 0 face of the three lowest layers)."""
import sys

import numpy as np
import pytest

from msg123_amd import gpu
from oracle import orc
from tests.golden.cases import CASES

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _ctx(z, **kw):
    box = gpu.BoxMesh(CASES["bnd"]())
    ctx = gpu.Context(box, **kw)
    bt = {k[3:]: v for k, v in z.items() if k.startswith("bt_")}
    ctx.set_bterms(bt)
    return box, ctx, bt


def test_bterms_first_assembly_vs_reference(golden):
    """calc_func and make_matvec at the first Newton iteration: F (= -rhs)
    within 1e-12 of the reference's largest entry (device vs glibc pow), the
    FD Jacobian within 1e-6 relative; the boundary terms move F and the
    diagonal by far more than that."""
    _say("test_bterms_first_assembly_vs_reference")
    z = golden("bnd")
    box, ctx, bt = _ctx(z)
    a = box.arrays()
    h0 = a["h0"]
    rech = a["read_rech"].astype(np.float64) * a["rech_area"]
    ctx.set_state(head_new=h0, head_old=h0, srat_old=z["a1_srat_old"], surf_head=h0[:a["ns"]],
                  surf_rati=np.zeros(a["ns"]), calc_rech=rech, delt=float(z["a1_delt"]))
    scale = np.abs(z["a1_rhs"]).max()
    f = ctx.calc_func(h0)
    np.testing.assert_allclose(-f, z["a1_rhs"], rtol=1e-12, atol=1e-12 * scale)
    rhs, dmat, lumat = ctx.make_matvec()
    np.testing.assert_allclose(rhs, z["a1_rhs"], rtol=1e-12, atol=1e-12 * scale)
    np.testing.assert_allclose(dmat, z["a1_dmat"], rtol=1e-6)
    np.testing.assert_allclose(lumat, z["a1_lumat"], rtol=1e-6,
                               atol=1e-6 * np.abs(z["a1_lumat"]).max())
    # the same state without the terms is far from the reference's
    ctx.set_bterms(None)
    rhs0, dmat0, _ = ctx.make_matvec()
    cells = np.unique(np.concatenate([bt["rive_cell"], bt["lake_cell"], bt["well_conn"],
                                      bt["seal_cell"]]))
    moved = (np.abs(rhs0 - z["a1_rhs"]) > 1e-9 * scale) | \
        (np.abs(dmat0 - z["a1_dmat"]) > 1e-6 * np.abs(z["a1_dmat"]))
    assert np.all(moved[cells])
    ctx.close()


def test_bterms_wet_state_vs_reference(golden):
    """make_matvec at the reference's ponded state (every head 2 cm above the
    top surface: every river and lake exchanges with the cell head):
    saturated everywhere, so only summation order may differ (1e-12)."""
    _say("test_bterms_wet_state_vs_reference")
    z = golden("bnd")
    box, ctx, _ = _ctx(z)
    a = box.arrays()
    prev = z["step_heads"][-2]
    so, _, _ = ctx.calc_srat_rperm(prev)
    rech = a["read_rech"].astype(np.float64) * a["rech_area"]
    ctx.set_state(head_new=z["wet_head"], head_old=prev, srat_old=so,
                  surf_head=z["wet_surf_head"], surf_rati=np.zeros(a["ns"]), calc_rech=rech,
                  delt=float(z["step_dts"][-1]))
    rhs, dmat, lumat = ctx.make_matvec()
    np.testing.assert_allclose(rhs, z["wet_rhs"], rtol=1e-12,
                               atol=1e-12 * np.abs(z["wet_rhs"]).max())
    np.testing.assert_allclose(dmat, z["wet_dmat"], rtol=1e-12)
    np.testing.assert_allclose(lumat, z["wet_lumat"], rtol=1e-12, atol=1e-300)
    ctx.close()


def test_bterms_calc_vecjacf_vs_oracle(golden):
    """calc_vecjacf with the terms (the well and sea cells enter F after the
    fused kernel, so J.v takes the separate path) against the oracle's."""
    _say("test_bterms_calc_vecjacf_vs_oracle")
    z = golden("bnd")
    box, ctx, bt = _ctx(z)
    a = box.arrays()
    h0 = a["h0"]
    rech = a["read_rech"].astype(np.float64) * a["rech_area"]
    ctx.set_state(head_new=h0, head_old=h0, srat_old=z["a1_srat_old"], surf_head=h0[:a["ns"]],
                  surf_rati=np.zeros(a["ns"]), calc_rech=rech, delt=float(z["a1_delt"]))
    rng = np.random.default_rng(9)
    x = h0 + rng.uniform(-0.02, 0.0, a["n"])
    v = rng.standard_normal(a["n"]) * 1e-3
    jv = ctx.calc_vecjacf(x, v)
    om = orc.OracleMesh(CASES["bnd"]())
    om.set_bterms(bt)
    jo = orc.OracleState(om).first_step().calc_vecjacf(x, v)
    np.testing.assert_allclose(jv, jo, rtol=0, atol=1e-8 * np.abs(jo).max())
    ctx.close()


@pytest.mark.parametrize("path", ["plain", "mixed"])
def test_bterms_time_loop_vs_reference(path, golden):
    """The time loop with the terms: the plain path (reference algorithm)
    under the reference's time control takes the same step times with Newton
    counts within one; the mixed path on the reference's step times; heads
    within 1e-6 m of the reference at every step."""
    _say("test_bterms_time_loop_vs_reference", path)
    z = golden("bnd")
    if path == "plain":
        box, ctx, _ = _ctx(z)
        res = ctx.run()
        np.testing.assert_array_equal(res["times"], z["step_times"])
        assert np.abs(res["newton"] - z["step_newton"]).max() <= 1
    else:
        box, ctx, _ = _ctx(z, precon=gpu.PRECON_LINE32, consistent_jacobian=True)
        res = ctx.run(schedule=z["step_times"])
        np.testing.assert_array_equal(res["times"], z["step_times"])
    err = np.abs(res["heads"] - z["step_heads"]).max() * z["len_scal"]
    _say("  max|dh| m", err, "newton", res["newton"].tolist(), "ref", z["step_newton"].tolist())
    assert err <= 1e-6
    ctx.close()


def test_bterms_mass_balance_vs_reference_file(golden):
    """The cumulative mass balance with the terms against the reference's own
    mass-balance file of the same run (whole domain, m^3, E13.6): flow,
    storage, recharge, well, river, lake, sea and the total, each to the
    file's six significant digits (or 1e-12 of the largest term where it
    prints less)."""
    _say("test_bterms_mass_balance_vs_reference_file")
    z = golden("bnd")
    box, ctx, _ = _ctx(z)
    ctx.set_mass_balance(True)
    res = ctx.run()
    assert res["nstep"] == len(z["step_times"])
    mb = ctx.mass_balance()
    L3 = float(np.float32(float(z["len_scal"]) ** 3))
    row = z["mass_values"][-1]
    assert np.float32(row[0]) == np.float32(z["step_times"][-1])
    ref = dict(connection=row[1], storage=row[2], recharge=row[3], well=row[4], surface=row[5],
               river=row[6], lake=row[7], sea=row[8], total=row[9])
    scale = float(np.abs(row[1:9]).max())
    for q, v in ref.items():
        got = mb[q] * L3
        half = 0.0 if v == 0 else 0.5 * 10.0 ** (np.floor(np.log10(abs(v))) + 1 - 6)
        # the total is the balance error of the Newton solution (5e-7 of
        # terms of 7 m^3): it moves with the last Newton iterate, here by
        # 2e-10 m^3 for heads 4e-11 m apart, so it is held to 1e-10 of the
        # largest term; a missing or mis-signed term moves it by 1e-2 and more
        tol = 1e-10 * scale if q == "total" else 1e-12 * scale
        _say(f"  {q}: {got:.6e} ref {v:.6e}")
        assert abs(got - v) <= 1.01 * half + tol, q
    for q in ("well", "river", "lake", "sea"):
        assert abs(ref[q]) > 1e-3 * scale, q  # the terms are really there
    ctx.close()


def test_bterms_rejected_where_they_do_not_fit(golden):
    """Out-of-range cells (a river off the top layer, a sea level index past
    read_seal) are MSG123_EINVAL; the noalloc mass balance refuses a context
    with terms; removing the terms restores the closed box's results."""
    _say("test_bterms_rejected_where_they_do_not_fit")
    z = golden("bnd")
    box, ctx, bt = _ctx(z)
    bad = dict(bt, rive_cell=bt["rive_cell"] + box.arrays()["ns"])
    with pytest.raises(RuntimeError):
        ctx.set_bterms(bad)
    bad = dict(bt, seal_sea=bt["seal_sea"] + 3)
    with pytest.raises(RuntimeError):
        ctx.set_bterms(bad)
    ctx.set_bterms(bt)
    with pytest.raises(RuntimeError):
        ctx.calc_mass_noalloc()
    ctx.set_bterms(None)
    plain = gpu.Context(gpu.BoxMesh(CASES["bnd"]()))
    r0 = plain.run(nstep_max=1)
    r1 = ctx.run(nstep_max=1)
    np.testing.assert_array_equal(r0["heads"], r1["heads"])
    ctx.close()
    plain.close()
