"""GPU parity of the top boundary's active seepage face and of the time
loop's failure budget.

The seepage face (func_surfterm, Calc_function.f90:304-391, and its
Jacobian in set_surfbound, Make_linearsystem.f90:410-497) is active where a
top cell's head reaches its surface bottom.  Two fixtures of the reference
build reach it: the fully saturated, ponded state the harness assembles after
each run (wet_*), and the `seep` run, whose top layer saturates cell by cell
during the time loop (its time loops run in tests/test_gpu_parity.py)."""
import sys

import numpy as np
import pytest

from msg123_amd import configs, gpu
from tests.golden.cases import CASES, FULL

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("consistent", [False, True])
@pytest.mark.parametrize("tag", FULL)
def test_wet_state_make_matvec(tag, consistent, golden):
    """make_matvec at the reference's ponded state (every head 2 cm above the
    top surface, surf_head the last converged top heads, old level the last
    accepted step) against the reference's rhs, dmat and lumat.  Every top
    face is active there, so the consistent Jacobian equals the reference's."""
    _say("test_wet_state_make_matvec", tag, consistent)
    z = golden(tag)
    box = gpu.BoxMesh(CASES[tag]())
    ctx = gpu.Context(box, consistent_jacobian=consistent)
    a = box.arrays()
    prev = z["step_heads"][-1]
    if len(z["step_heads"]) > 1:
        prev = z["step_heads"][-2]
    so, _, _ = ctx.calc_srat_rperm(prev)
    p = CASES[tag]()
    if p.rech_etime is None:
        rech = a["read_rech"].astype(np.float64) * a["rech_area"]
    else:
        # the last step's recharge: the last record's rate, converted in
        # float32 as assign_surfbv converts it (Assign_boundary.f90:216)
        f32 = np.float32
        rate = (f32(p.rech_mmh[-1]) * (f32(1.0) / f32(a["len_scal"]))) * (f32(1.0e-3) / f32(3600.0))
        rech = np.float64(rate) * a["rech_area"]
    ctx.set_state(head_new=z["wet_head"], head_old=prev, srat_old=so,
                  surf_head=z["wet_surf_head"], surf_rati=np.zeros(a["ns"]), calc_rech=rech,
                  delt=float(z["step_dts"][-1]))
    rhs, dmat, lumat = ctx.make_matvec()
    assert np.all(z["wet_head"] >= a["surf_bott"][0])
    # saturated everywhere: no pow, so only summation order can differ
    np.testing.assert_allclose(rhs, z["wet_rhs"], rtol=1e-12, atol=1e-12 * np.abs(z["wet_rhs"]).max())
    np.testing.assert_allclose(dmat, z["wet_dmat"], rtol=1e-12)
    np.testing.assert_allclose(lumat, z["wet_lumat"], rtol=1e-12, atol=1e-300)
    st = ctx.get_state()
    # the active face sets surf_rati = 1 (Calc_function.f90:330-360)
    assert np.all(st["surf_rati"] == 1.0)


def test_seep_fixture_saturates_the_top(golden):
    """The seep fixture does what it is for: its top layer goes from no
    saturated cell to several during the run, so the time loops on it cross
    the seepage face's switch."""
    z = golden("seep")
    ns = int(z["nx"]) * int(z["ny"])
    bott = z["st_surf_bott"]
    sat = [(h[:ns] >= bott).sum() for h in z["step_heads"]]
    _say("  saturated top cells per step", sat)
    assert sat[0] == 0 and sat[-1] >= 3


def test_run_box_failure_budget():
    """A run that cannot converge (one Newton iteration per attempt) ends
    with MSG123_ECONV after max_fail consecutive failed attempts, or once
    max_newton Newton iterations went into failed attempts, instead of
    halving the step until dt < 1e-20 s; a normal run is unaffected."""
    _say("test_run_box_failure_budget")
    p = configs.small(6, 5, 4, seed=3, nsteps=3)
    box = gpu.BoxMesh(p)
    ctx = gpu.Context(box, maxout_iter=1)
    with pytest.raises(RuntimeError) as e:
        ctx.run(max_fail=4)
    assert e.value.code == gpu.ECONV
    with pytest.raises(RuntimeError) as e:
        ctx.run(max_newton=3)
    assert e.value.code == gpu.ECONV
    ok = gpu.Context(box)
    res = ok.run(max_fail=2)
    assert res["nstep"] == 3
