"""GPU parity: every entry point of libmsg123_gpu.so against the oracle and the
reference build's fixtures.  Tolerances: integer/index work and the D-ILU
sweeps bit-exact; residual and Jacobian within the device-vs-glibc pow
difference (stated per test); heads within 1e-6 m with Newton counts equal."""
import os
import sys

import numpy as np
import pytest

from msg123_amd import configs, gpu
from oracle import orc
from tests.golden.cases import AMG, CASES, FULL

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _ctx(p, **kw):
    box = gpu.BoxMesh(p)
    return box, gpu.Context(box, **kw)


def _first_step_state(ctx, box, z):
    a = box.arrays()
    h0 = a["h0"]
    so = z["a1_srat_old"]
    rech = a["read_rech"].astype(np.float64) * a["rech_area"]
    ctx.set_state(head_new=h0, head_old=h0, srat_old=so, surf_head=h0[:a["ns"]],
                  surf_rati=np.zeros(a["ns"]), calc_rech=rech, delt=float(z["a1_delt"]))
    return a


@pytest.mark.parametrize("tag", FULL)
def test_srat_rperm(tag, golden):
    _say("test_srat_rperm", tag)
    z = golden(tag)
    box, ctx = _ctx(CASES[tag]())
    a = box.arrays()
    rng = np.random.default_rng(5)
    h = a["h0"] + rng.uniform(-0.3, 0.3, a["n"])
    s, k, q = ctx.calc_srat_rperm(h)
    om = orc.OracleMesh(CASES[tag]())
    so, ko, qo = np.zeros(a["n"]), np.zeros(a["n"]), np.zeros(a["n"])
    orc.lib().orc_srat_rperm(om.m, 0.0, *[x.ctypes.data_as(orc._dp) for x in (h, so, ko, qo)])
    # device pow vs glibc pow: a few ulp
    np.testing.assert_allclose(s, so, rtol=1e-14, atol=0)
    np.testing.assert_allclose(k, ko, rtol=1e-12, atol=1e-300)
    np.testing.assert_array_equal(q, qo)
    del z


@pytest.mark.parametrize("tag", FULL)
def test_first_assembly(tag, golden):
    """make_matvec at the first Newton iteration vs the reference build."""
    _say("test_first_assembly", tag)
    z = golden(tag)
    box, ctx = _ctx(CASES[tag]())
    _first_step_state(ctx, box, z)
    f = ctx.calc_func(z["a1_head_new"])
    np.testing.assert_allclose(-f, z["a1_rhs"], rtol=1e-12, atol=1e-12 * np.abs(z["a1_rhs"]).max())
    rhs, dmat, lumat = ctx.make_matvec()
    np.testing.assert_allclose(rhs, z["a1_rhs"], rtol=1e-12, atol=1e-12 * np.abs(z["a1_rhs"]).max())
    # FD derivatives with eps ~ 1e-8 amplify pow's last-ulp differences
    np.testing.assert_allclose(dmat, z["a1_dmat"], rtol=1e-6)
    np.testing.assert_allclose(lumat, z["a1_lumat"], rtol=1e-6, atol=1e-6 * np.abs(z["a1_lumat"]).max())
    st = ctx.get_state()
    np.testing.assert_allclose(st["srat_new"], z["a1_srat_new"], rtol=1e-14)
    np.testing.assert_allclose(st["rel_perm"], z["a1_rel_perm"], rtol=1e-12)


@pytest.mark.parametrize("tag", FULL)
def test_dilu_sweeps_bit_exact(tag, golden):
    """precon_dilu, solve_ilu and calc_matvec on the reference's own system:
    bit-identical to the serial sweeps of the oracle."""
    _say("test_dilu_sweeps_bit_exact", tag)
    z = golden(tag)
    box, ctx = _ctx(CASES[tag]())
    om = orc.OracleMesh(CASES[tag]())
    ctx.load_system(z["a1_rhs"], z["a1_dmat"], z["a1_lumat"])
    d = ctx.precon_dilu()
    d_ref = orc.precon_dilu(om, z["a1_dmat"], z["a1_lumat"])
    assert np.array_equal(d, d_ref)
    r = np.random.default_rng(11).standard_normal(len(d))
    zz = ctx.solve_ilu(r)
    assert np.array_equal(zz, orc.solve_ilu(om, r, d_ref, z["a1_lumat"]))
    y = ctx.matvec(r)
    assert np.array_equal(y, orc.matvec(om, z["a1_dmat"], z["a1_lumat"], r))


@pytest.mark.parametrize("launch", [gpu.ILU_WAVE, gpu.ILU_DIAGONALS])
def test_dilu_sweeps_bit_exact_large(launch):
    """Sweeps over many tiles and partial tiles (box 70x37x19) stay bit-exact,
    in one launch per sweep (tiles awaiting their predecessors' flags) and in
    one launch per tile diagonal; repeated sweeps (new epochs) too."""
    _say("test_dilu_sweeps_bit_exact_large", launch)
    p = configs.small(70, 37, 19, seed=23)
    box, ctx = _ctx(p)
    ctx.set_ilu_launch(launch)
    om = orc.OracleMesh(p)
    st = orc.OracleState(om).first_step()
    rhs, dmat, lumat = st.make_matvec()
    ctx.load_system(rhs, dmat, lumat)
    d = ctx.precon_dilu()
    d_ref = orc.precon_dilu(om, dmat, lumat)
    assert np.array_equal(d, d_ref)
    for seed in (2, 3, 4):
        r = np.random.default_rng(seed).standard_normal(len(d))
        assert np.array_equal(ctx.solve_ilu(r), orc.solve_ilu(om, r, d_ref, lumat))
    assert np.array_equal(ctx.matvec(r), orc.matvec(om, dmat, lumat, r))
    assert np.array_equal(ctx.precon_dilu(), d_ref)


@pytest.mark.parametrize("tag", FULL)
def test_solve_linalg(tag, golden):
    """BiCGStab on the reference system: same iteration count and convergence flag
    as the reference, solution within 1e-9 relative of the reference's."""
    _say("test_solve_linalg", tag)
    z = golden(tag)
    box, ctx = _ctx(CASES[tag]())
    ctx.load_system(z["a1_rhs"], z["a1_dmat"], z["a1_lumat"])
    dh, it, last, conv = ctx.solve_linalg(float(z["a1_bnorm"]), 0.5)
    assert it == int(z["a1_in_iter"]) and conv == int(z["a1_conv_flag"])
    scale = np.abs(z["a1_dh"]).max()
    assert np.abs(dh - z["a1_dh"]).max() <= 1e-9 * scale
    # tight tolerance: converge to numpy's dense solution
    om = orc.OracleMesh(CASES[tag]())
    x, it2, _, conv2 = orc.bicgstab(om, z["a1_dmat"], z["a1_lumat"], z["a1_rhs"],
                                    float(z["a1_bnorm"]), 1e-12, 1000)
    dh2, it3, _, conv3 = ctx.solve_linalg(float(z["a1_bnorm"]), 1e-12)
    assert conv3 == 1 and abs(it3 - it2) <= 1
    assert np.abs(dh2 - x).max() <= 1e-8 * np.abs(x).max()


@pytest.mark.parametrize("tag", FULL)
def test_time_loop_vs_reference(tag, golden):
    """Full run through msg123_run_box vs the reference's per-step FP64 heads."""
    _say("test_time_loop_vs_reference", tag)
    z = golden(tag)
    box, ctx = _ctx(CASES[tag]())
    res = ctx.run()
    assert res["nstep"] == len(z["step_times"])
    np.testing.assert_array_equal(res["times"], z["step_times"])
    assert np.abs(res["newton"] - z["step_newton"]).max() <= 1
    err = np.abs(res["heads"] - z["step_heads"]).max() * z["len_scal"]
    _say("  max|dh| m", err, "newton", res["newton"].tolist())
    assert err <= 1e-6


@pytest.mark.parametrize("tag", ["box12", "seep"])
def test_calc_vecjacf_vs_oracle(tag, golden):
    """msg123_calc_vecjacf (Calc_function.f90:515-621) against the oracle's
    at the same perturbed state, and against the assembled Jacobian:
    -(F(x + eps v) - F(x))/eps = -J v to the difference quotients' accuracy
    (J the FD Jacobian of make_matvec, dF/dh: J dh = rhs = -F)."""
    _say("test_calc_vecjacf_vs_oracle", tag)
    z = golden(tag)
    p = CASES[tag]()
    box, ctx = _ctx(p)
    a = _first_step_state(ctx, box, z)
    rng = np.random.default_rng(7)
    x = a["h0"] + rng.uniform(-0.05, 0.0, a["n"])
    v = rng.standard_normal(a["n"]) * 1e-3
    jv = ctx.calc_vecjacf(x, v)
    st = orc.OracleState(orc.OracleMesh(p)).first_step()
    jo = st.calc_vecjacf(x, v)
    # the step's dots are summed in another order (eps differs in the last
    # bits) and device pow differs from glibc pow by ulps, both divided by eps
    np.testing.assert_allclose(jv, jo, rtol=0, atol=1e-8 * np.abs(jo).max())
    ctx.set_state(head_new=x)
    rhs, dmat, lumat = ctx.make_matvec()
    ctx.load_system(rhs, dmat, lumat)
    av = ctx.matvec(v)
    rel = np.linalg.norm(jv + av) / np.linalg.norm(av)
    _say("  |jv - jv_oracle| / |jv| =", np.abs(jv - jo).max() / np.abs(jo).max(),
         " |jv + Av| / |Av| =", rel)
    assert rel <= 3e-2


@pytest.mark.parametrize("control", ["device", "graph"])
@pytest.mark.parametrize("path", ["plain", "mixed", "amg_ref"])
@pytest.mark.parametrize("tag", ["box9", "seep"])
def test_device_control_matches_host_control(tag, path, control, golden):
    """msg123_set_control: the device-driven Newton iterations (the Krylov and
    Newton scalars in device memory, one host wait per iteration), enqueued
    kernel by kernel or replayed from captured HIP graphs, take the same
    steps as the host-driven ones: same Newton counts, heads equal to 1e-12 m
    (set_eater's pow on the device may differ from the host's in the last
    bit).  Under graph control every step replays the graphs again, with the
    time loop's own kernels and synchronisations between the replays."""
    _say("test_device_control_matches_host_control", tag, path, control)
    z = golden(tag)
    runs = []
    for mode in (gpu.CONTROL_HOST,
                 gpu.CONTROL_DEVICE if control == "device" else gpu.CONTROL_GRAPH):
        if path == "mixed":
            box, ctx = _ctx(CASES[tag](), precon=gpu.PRECON_LINE32, consistent_jacobian=True)
            ctx.set_law(gpu.LAW_LOGEXP)
            ctx.set_newton(**gpu.MIXED_NEWTON)
        else:
            box, ctx = _ctx(CASES[tag](), precon=gpu.PRECON_AMG if path == "amg_ref"
                            else gpu.PRECON_DILU)
        ctx.set_control(mode)
        runs.append(ctx.run(schedule=z["step_times"]))
        ctx.close()
    a, b = runs
    _say("  newton host", a["newton"].tolist(), "device", b["newton"].tolist())
    np.testing.assert_array_equal(a["times"], b["times"])
    assert np.array_equal(a["newton"], b["newton"])
    assert np.abs(a["heads"] - b["heads"]).max() * z["len_scal"] <= 1e-12


def test_config0_vs_reference(golden):
    """Config 0 (40x40x20, 24 h): step times, Newton counts (+-1), heads within 1e-6 m."""
    _say("test_config0_vs_reference")
    z = golden("c0")
    box, ctx = _ctx(configs.config0())
    res = ctx.run()
    np.testing.assert_array_equal(res["times"], z["step_times"])
    assert np.abs(res["newton"] - z["step_newton"]).max() <= 1
    err = np.abs(res["heads"][:, z["sample_idx"]] - z["step_heads_sample"]).max() * z["len_scal"]
    _say("  max|dh| m", err, "newton", res["newton"].tolist())
    assert err <= 1e-6


def test_mass_balance_closed_box():
    """Closed box with top recharge: sum of storage change equals recharge
    volume over each converged step (the flux terms cancel pairwise), to the
    Newton tolerance."""
    _say("test_mass_balance_closed_box")
    p = configs.small(10, 9, 8, seed=31, nsteps=3)
    box, ctx = _ctx(p)
    res = ctx.run()
    a = box.arrays()
    om = orc.OracleMesh(p)
    st = orc.OracleState(om).first_step()
    prev = a["h0"]
    total_in = 0.0
    total_err = 0.0
    for k in range(res["nstep"]):
        h = res["heads"][k]
        dt = res["dts"][k]
        s1, _, q1 = _vg(om, h)
        s0, _, q0 = _vg(om, prev)
        poro = a["poro"].astype(np.float64)
        # storage with the current-iterate Ss switch for both levels (Calc_function.f90:53)
        dstore = ((q1 * s1 * h - q1 * s0 * prev) + poro * (s1 - s0)) * a["cell_vol"]
        rech = (a["read_rech"].astype(np.float64) * a["rech_area"]).sum() * dt
        total_in += rech
        total_err += dstore.sum() - rech
        prev = h
    assert abs(total_err) <= 1e-6 * total_in, (total_err, total_in)
    del st


@pytest.mark.parametrize("tag", ["box9", "seep", "rech"])
def test_mass_balance_vs_reference_file(tag):
    """msg123_calc_mass accumulated over the reference's output windows
    against the reference's own mass-balance file of the same run
    (write_out_massf, Write_output.f90:497-538, whole domain, m^3, E13.6):
    flow, storage, recharge, surface and the sum of all terms (the balance
    error, five to seven orders below the terms) each to the file's six
    significant digits, or 1e-12 of the run's largest window of recharge
    where the file prints a value below that."""
    _say("test_mass_balance_vs_reference_file", tag)
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_mass.npz"))
    vals, times = z[f"{tag}_values"], z[f"{tag}_step_times"]
    L3 = float(np.float32(float(z[f"{tag}_len_scal"]) ** 3))
    box, ctx = _ctx(CASES[tag]())
    ctx.set_mass_balance(True)
    prev = dict(storage=0.0, connection=0.0, recharge=0.0, surface=0.0, total=0.0)
    # the run's scale: its largest window of recharge (the flows cancel
    # pairwise in the closed box, so FLOW is summation noise on both sides)
    scale = float(np.abs(vals[:, [2, 3]]).max())
    done = 0
    for row in vals:
        k = int(np.searchsorted(times, row[0], side="right")) - done
        assert k > 0 and np.float32(times[done + k - 1]) == np.float32(row[0])
        res = ctx.run(nstep_max=k, keep_heads=False, reset=(done == 0))
        assert res["nstep"] == k
        done += k
        mb = ctx.mass_balance()
        win = {q: (mb[q] - prev[q]) * L3 for q in prev}
        prev = {q: mb[q] for q in prev}
        ref = dict(connection=row[1], storage=row[2], recharge=row[3], surface=row[5],
                   total=row[9])
        _say("  t", row[0], {q: (win[q], ref[q]) for q in ref})
        for q in ("connection", "storage", "recharge", "surface", "total"):
            # half a unit of the sixth significant digit the file prints
            # (0.dddddd E+xx), plus the sums' own rounding
            v = ref[q]
            half = 0.0 if v == 0 else 0.5 * 10.0 ** (np.floor(np.log10(abs(v))) + 1 - 6)
            assert abs(win[q] - v) <= 1.01 * half + 1e-12 * scale, q
    ctx.close()


def test_mass_balance_accounting_vs_oracle():
    """msg123_calc_mass accumulated over the accepted steps (calc_mass,
    Calc_function.f90:103-174, domain sums as calc_out_mass) against a numpy
    restatement on the oracle's constitutive law from the returned heads:
    storage (old and new Ss switch), connection flow (cancels pairwise),
    recharge, and the Newton residual at the accepted heads."""
    _say("test_mass_balance_accounting_vs_oracle")
    p = configs.small(10, 9, 8, seed=31, nsteps=4)
    box, ctx = _ctx(p)
    ctx.set_mass_balance(True)
    res = ctx.run()
    mb = ctx.mass_balance()
    a = box.arrays()
    om = orc.OracleMesh(p)
    n = a["n"]
    rows = np.repeat(np.arange(n), np.diff(a["offind"]))
    cols = a["offrow"]
    poro = a["poro"].astype(np.float64)
    rech = a["read_rech"].astype(np.float64) * a["rech_area"]
    prev = a["h0"]
    s_prev = _vg(om, prev)[0]
    want = dict(storage=0.0, connection=0.0, recharge=0.0, residual=0.0)
    for k in range(res["nstep"]):
        h, dt = res["heads"][k], res["dts"][k]
        s1, k1, q1 = _vg(om, h)
        _, _, q0 = _vg(om, prev)
        stom = -((q1 * s1 * h - q0 * s_prev * prev) + poro * (s1 - s_prev)) * a["cell_vol"] / dt
        stof = -((q1 * s1 * h - q1 * s_prev * prev) + poro * (s1 - s_prev)) * a["cell_vol"] / dt
        dhd = h[cols] - h[rows]
        kup = np.where(-dhd >= 0.0, k1[rows], k1[cols])
        flow = np.bincount(rows, weights=a["hydf"] * kup * a["abyd"] * dhd, minlength=n)
        f = stof + flow
        f[:a["ns"]] += rech
        want["storage"] += stom.sum() * dt
        want["connection"] += flow.sum() * dt
        want["recharge"] += rech.sum() * dt
        want["residual"] += f.sum() * dt
        prev, s_prev = h, s1
    scale = mb["recharge_abs"]
    _say("  device", mb, "oracle", want)
    assert mb["steps"] == res["nstep"]
    assert abs(mb["recharge"] - want["recharge"]) <= 1e-12 * scale
    # srat_old carries the saturation of the last assembly, within the
    # Newton tolerance of S(h_prev)
    assert abs(mb["storage"] - want["storage"]) <= 1e-7 * scale
    assert abs(mb["connection"]) <= 1e-10 * scale
    assert abs(mb["residual"] - want["residual"]) <= 1e-7 * scale
    # the storage change balances the recharge to the Newton tolerance
    assert abs(mb["total"]) <= 1e-6 * scale
    assert abs(mb["residual"]) <= 1e-6 * scale


def _vg(om, h):
    n = om.m.n
    h = np.ascontiguousarray(h)
    s, k, q = np.zeros(n), np.zeros(n), np.zeros(n)
    orc.lib().orc_srat_rperm(om.m, 0.0, *[x.ctypes.data_as(orc._dp) for x in (h, s, k, q)])
    return s, k, q


# ---------------------------------------------------------------------------
# mixed-precision path: FP32 vertical-line preconditioner, FP64 Krylov and Newton

def test_line32_preconditioner_is_column_tridiagonal_solve():
    """M^-1 r of the FP32 line preconditioner equals exact FP64 tridiagonal
    solves of every vertical column to FP32 accuracy."""
    _say("test_line32_preconditioner_is_column_tridiagonal_solve")
    from scipy.linalg import solve_banded
    p = configs.small(9, 8, 12, seed=41)
    box, ctx = _ctx(p, precon=gpu.PRECON_LINE32)
    om = orc.OracleMesh(p)
    st = orc.OracleState(om).first_step()
    st.set_head(st.vec("head_new") + np.linspace(-0.03, 0.03, om.m.n))
    rhs, dmat, lumat = st.make_matvec()
    ctx.load_system(rhs, dmat, lumat)
    ctx.precon_dilu()
    r = np.random.default_rng(4).standard_normal(om.m.n)
    z = ctx.solve_ilu(r)
    s = om.static()
    nxy, nz = p.nx * p.ny, p.nz
    up = np.zeros(om.m.n)
    dn = np.zeros(om.m.n)
    for i in range(om.m.n):
        for k in range(s["offind"][i], s["offind"][i + 1]):
            if s["dir"][k] == 1:
                up[i] = lumat[k]
            elif s["dir"][k] == 6:
                dn[i] = lumat[k]
    zref = np.zeros(om.m.n)
    for t in range(nxy):
        idx = t + nxy * np.arange(nz)
        ab = np.zeros((3, nz))
        ab[0, 1:] = dn[idx[:-1]]
        ab[1] = dmat[idx]
        ab[2, :-1] = up[idx[1:]]
        zref[idx] = solve_banded((1, 1), ab, r[idx])
    assert np.abs(z - zref).max() <= 1e-4 * np.abs(zref).max()


@pytest.mark.parametrize("dims", [(9, 8, 12), (7, 5, 3)])
def test_mcilu32_factor_and_apply_bit_exact(dims):
    """FP32 multicolour ILU(0): the FP64 pivots and z = M^-1 r equal the
    oracle's (oracle/mcilu.py, FP32 passes in the device's operation order)
    bit for bit, on an FD Jacobian of a small box (an odd cell count too)."""
    from oracle import mcilu
    _say("test_mcilu32_factor_and_apply_bit_exact", dims)
    nx, ny, nz = dims
    p = configs.small(nx, ny, nz, seed=43)
    box, ctx = _ctx(p, precon=gpu.PRECON_MCILU32)
    om = orc.OracleMesh(p)
    st = orc.OracleState(om).first_step()
    st.set_head(st.vec("head_new") + np.linspace(-0.03, 0.03, om.m.n))
    rhs, dmat, lumat = st.make_matvec()
    ctx.load_system(rhs, dmat, lumat)
    d = ctx.precon_dilu()
    s = om.static()
    A7 = mcilu.from_csr(s["offind"], s["offrow"], s["dir"], dmat, lumat)
    piv = mcilu.factor(A7, nx, ny, nz)
    np.testing.assert_array_equal(d, piv)
    for seed in (4, 5):
        r = np.random.default_rng(seed).standard_normal(om.m.n)
        z = ctx.solve_ilu(r)
        np.testing.assert_array_equal(z, mcilu.apply(A7, piv, r, nx, ny, nz, fp32=True))
    ctx.close()


@pytest.mark.parametrize("tag", FULL)
def test_mcilu32_time_loop_vs_reference(tag, golden):
    """BiCGStab with the FP32 multicolour ILU(0) replaying the reference's
    step times reaches the same FP64 Newton tolerance: heads within 1e-6 m
    of the reference at every step."""
    _say("test_mcilu32_time_loop_vs_reference", tag)
    z = golden(tag)
    box, ctx = _ctx(CASES[tag](), precon=gpu.PRECON_MCILU32, consistent_jacobian=True)
    res = ctx.run(schedule=z["step_times"])
    np.testing.assert_array_equal(res["times"], z["step_times"])
    err = np.abs(res["heads"] - z["step_heads"]).max() * z["len_scal"]
    _say("  max|dh| m", err, "newton", res["newton"].tolist(), "ref", z["step_newton"].tolist())
    assert err <= 1e-6
    ctx.close()


def test_logexp_law_vs_extended_precision():
    """MSG123_LAW_LOGEXP against the closed form (van Genuchten 1980, Mualem
    1976) in extended precision, written in its stable form
    1 - (1 - Se^(1/m))^m = -expm1(-m log1p(1/beta)): S within 1e-14 and kr
    within 1e-13 relative over dry to saturated heads (the reference's pow
    form is off by up to 4e-10 in dry cells by cancellation)."""
    _say("test_logexp_law_vs_extended_precision")
    p = configs.small(9, 8, 7, seed=43)
    box, ctx = _ctx(p)
    ctx.set_law(gpu.LAW_LOGEXP)
    a = box.arrays()
    n = a["n"]
    rng = np.random.default_rng(8)
    psi = -np.exp(rng.uniform(np.log(1e-7), np.log(30.0), n)) / a["len_scal"]
    psi[::17] = 0.0
    psi[::19] = 1e-3
    h = a["cell_top"] + psi
    s, k, q = ctx.calc_srat_rperm(h)
    L = np.longdouble
    ph = (L(h) - L(a["cell_top"])) * L(a["len_scal"]) / L(a["len_scal"])
    x = np.abs(ph * L(a["reta"].astype(np.float64)))
    vn = L(a["retn"].astype(np.float64))
    m = 1 - 1 / vn
    with np.errstate(divide="ignore", over="ignore"):
        b = x ** vn
        se = (1 + b) ** (-m)
        c = -np.expm1(-m * np.log1p(1 / b))
    poro = a["poro"].astype(np.float64)
    resi = a["resi"].astype(np.float64)
    dsp = (a["poro"] - a["resi"]).astype(np.float32).astype(np.float64)
    s_ref = np.where(ph < 0, (L(resi) + L(dsp) * se) / L(poro), 1.0).astype(np.float64)
    k_ref = np.where(ph < 0, np.sqrt(se) * c * c, 1.0).astype(np.float64)
    np.testing.assert_allclose(s, s_ref, rtol=1e-14, atol=0)
    np.testing.assert_allclose(k, k_ref, rtol=1e-13, atol=1e-300)
    np.testing.assert_array_equal(q, np.where(ph < 0, 0.0, a["ss"].astype(np.float64)))


def test_logexp_jacobian_from_analytic_derivative():
    """The log/exp law's Jacobian takes kr(h + newper) and S(h + newper) from
    the law's analytic derivative (vg_pair): at an unsaturated state its
    make_matvec equals the reference law's, whose second law evaluation
    gives the difference quotients, to their accuracy (1e-6 of the largest
    entry; a wrong factor in a derivative is off by 1e-1 and more), and its
    rhs to the two laws' difference (1e-9)."""
    _say("test_logexp_jacobian_from_analytic_derivative")
    p = configs.small(9, 8, 7, seed=43)
    om = orc.OracleMesh(p)
    st = orc.OracleState(om).first_step()
    a = gpu.BoxMesh(p).arrays()
    rng = np.random.default_rng(12)
    h = a["h0"] - rng.uniform(0.05, 0.4, a["n"]) / a["len_scal"]
    out = []
    for law in (gpu.LAW_REFERENCE, gpu.LAW_LOGEXP):
        box, ctx = _ctx(p)
        ctx.set_law(law)
        ctx.set_state(head_new=h, head_old=a["h0"], srat_old=st.vec("srat_old"),
                      surf_head=a["h0"][:a["ns"]], surf_rati=np.zeros(a["ns"]),
                      calc_rech=a["read_rech"].astype(np.float64) * a["rech_area"],
                      delt=float(st.s.delt))
        out.append(ctx.make_matvec())
        ctx.close()
    (r0, d0, l0), (r1, d1, l1) = out
    assert np.all(h - a["cell_top"] < 0)
    np.testing.assert_allclose(r1, r0, rtol=0, atol=1e-9 * np.abs(r0).max())
    err_d = np.abs(d1 - d0).max() / np.abs(d0).max()
    err_l = np.abs(l1 - l0).max() / np.abs(l0).max()
    _say("  dmat", err_d, "lumat", err_l)
    assert err_d <= 1e-6 and err_l <= 1e-6


@pytest.mark.parametrize("beta_refine", [True, False])
@pytest.mark.parametrize("tag", FULL)
def test_logexp_line32_time_loop_vs_reference(tag, beta_refine, golden):
    """The mixed path (FP32 line preconditioner, consistent Jacobian,
    log/exp law, its forcing cap, algebraic slope; the bench runs it without
    the line search's beta refinement) on the reference's step times: heads
    within 1e-6 m of the reference."""
    _say("test_logexp_line32_time_loop_vs_reference", tag, beta_refine)
    z = golden(tag)
    box, ctx = _ctx(CASES[tag](), precon=gpu.PRECON_LINE32, consistent_jacobian=True)
    ctx.set_law(gpu.LAW_LOGEXP)
    ctx.set_newton(**dict(gpu.MIXED_NEWTON, beta_refine=beta_refine))
    res = ctx.run(schedule=z["step_times"])
    np.testing.assert_array_equal(res["times"], z["step_times"])
    err = np.abs(res["heads"] - z["step_heads"]).max() * z["len_scal"]
    _say("  max|dh| m", err, "newton", res["newton"].tolist(), "ref", z["step_newton"].tolist())
    assert err <= 1e-6


@pytest.mark.parametrize("tag", FULL)
def test_line32_time_loop_vs_reference(tag, golden):
    """The mixed path (FP32 line preconditioner, Jacobian consistent with the
    residual) replaying the reference's step times reaches the same FP64
    Newton tolerance: heads within 1e-6 m of the reference at every step."""
    consistent = True
    _say("test_line32_time_loop_vs_reference", tag)
    z = golden(tag)
    box, ctx = _ctx(CASES[tag](), precon=gpu.PRECON_LINE32, consistent_jacobian=consistent)
    res = ctx.run(schedule=z["step_times"])
    np.testing.assert_array_equal(res["times"], z["step_times"])
    err = np.abs(res["heads"] - z["step_heads"]).max() * z["len_scal"]
    _say("  max|dh| m", err, "newton", res["newton"].tolist(), "ref", z["step_newton"].tolist())
    assert err <= 1e-6


def test_line32_config0_vs_reference(golden):
    """Config 0 on the reference's step times: heads within 1e-6 m; the
    consistent Jacobian removes most Newton iterations (466 -> about 100)."""
    consistent = True
    _say("test_line32_config0_vs_reference")
    z = golden("c0")
    box, ctx = _ctx(configs.config0(), precon=gpu.PRECON_LINE32, consistent_jacobian=consistent)
    res = ctx.run(schedule=z["step_times"])
    np.testing.assert_array_equal(res["times"], z["step_times"])
    err = np.abs(res["heads"][:, z["sample_idx"]] - z["step_heads_sample"]).max() * z["len_scal"]
    _say("  max|dh| m", err, "newton", res["newton"].sum(), "ref", z["step_newton"].sum())
    assert err <= 1e-6
    assert res["newton"].sum() < z["step_newton"].sum()


def test_schedule_replay_plain_path_is_identical(golden):
    """Replaying the reference's own step times through the schedule option gives
    the same heads as the heuristic time control (same steps, same solves)."""
    _say("test_schedule_replay_plain_path_is_identical")
    z = golden("box9")
    box, ctx = _ctx(CASES["box9"]())
    a = ctx.run()
    b = ctx.run(schedule=z["step_times"])
    np.testing.assert_array_equal(a["times"], b["times"])
    assert np.array_equal(a["heads"], b["heads"])


@pytest.mark.parametrize("tag", ["amg12", "amg_box9"])
def test_amg_path_time_loop_vs_reference(tag, golden):
    """The reference's AMG path (precon_type = 1: D-ILU numerics, slope and
    log from the level-1 system the V-cycle overwrote) against the
    reference's precon_type = 1 runs: step times, Newton counts (+-1) and
    heads within 1e-6 m."""
    _say("test_amg_path_time_loop_vs_reference", tag)
    z = golden(tag)
    box, ctx = _ctx(AMG[tag][0](), precon=gpu.PRECON_AMG)
    res = ctx.run()
    assert res["nstep"] == len(z["step_times"])
    np.testing.assert_array_equal(res["times"], z["step_times"])
    assert np.abs(res["newton"] - z["step_newton"]).max() <= 1
    err = np.abs(res["heads"] - z["step_heads"]).max() * z["len_scal"]
    _say("  max|dh| m", err, "newton", res["newton"].tolist())
    assert err <= 1e-6


def test_amg_path_quirk_changes_the_trajectory(golden):
    """Two BiCGStab iterations per solve (box9_m2 / amg_m2): the reference's
    AMG path takes the line search's slope from the last preconditioner input
    the V-cycle left in the level-1 rhs (Linear_solution.f90:525-527,
    Nonlinear_solution.f90:500), so its run parts from the D-ILU path's.
    On the GPU both paths follow their reference runs over the steps the two
    runs share (the first four: step times, Newton counts within 1, heads
    within 1e-6 m); later steps hinge on line searches at 1e-8 m changes,
    where the order of a dot product's partial sums already decides whether a
    step converges in 20 Newton iterations, so they are not compared.  The GPU
    AMG path runs the D-ILU path's kernels with the same rounding, so any
    difference between the two GPU runs is the quirk's."""
    _say("test_amg_path_quirk_changes_the_trajectory")
    runs = {}
    for tag, p, precon in (("amg_m2", AMG["amg_m2"][0](), gpu.PRECON_AMG),
                           ("box9_m2", CASES["box9_m2"](), gpu.PRECON_DILU)):
        z = golden(tag)
        box, ctx = _ctx(p, precon=precon)
        res = ctx.run()
        runs[tag] = res
        _say("  ", tag, "gpu", res["times"].tolist(), res["newton"].tolist(), "ref",
             z["step_times"].tolist(), z["step_newton"].tolist())
        np.testing.assert_array_equal(res["times"][:4], z["step_times"][:4])
        assert np.abs(res["newton"][:4] - z["step_newton"][:4]).max() <= 1
        err = np.abs(res["heads"][:4] - z["step_heads"][:4]).max() * z["len_scal"]
        _say("   max|dh| m over the first four steps", err)
        assert err <= 1e-6
    a, b = runs["amg_m2"], runs["box9_m2"]
    k = min(a["nstep"], b["nstep"])
    assert (a["nstep"] != b["nstep"] or not np.array_equal(a["times"], b["times"])
            or not np.array_equal(a["heads"][:k], b["heads"][:k]))


@pytest.mark.gpu
@pytest.mark.parametrize("precon,precon_type", [(gpu.PRECON_DILU, 0), (gpu.PRECON_AMG, 1)])
@pytest.mark.parametrize("k", [1, 3])
def test_iteration_log_vs_oracle(precon, precon_type, k, golden):
    """The convergence-log record of Newton iteration k (cell of the largest
    change, diagonal, right-hand side, head there) against the oracle's
    level-1 system after k iterations; on the AMG path that system holds the
    D-ILU diagonal and the last preconditioner input."""
    import ctypes as C
    _say("test_iteration_log_vs_oracle", precon, k)
    z = golden("box9")
    p = CASES["box9"]()
    box, ctx = _ctx(p, maxout_iter=k, precon=precon)
    _first_step_state(ctx, box, z)
    ctx.calc_numsol()
    recs, why = ctx.iter_log()
    assert len(recs) == k and why == gpu.END_MAXITER
    om = orc.OracleMesh(p)
    st = orc.OracleState(om).first_step()
    cfg = om.settings(precon_type)
    cfg.maxout_iter = k
    orc.lib().orc_calc_numsol(C.byref(om.m), C.byref(st.s), C.byref(cfg))
    L = z["len_scal"]
    h, hp = st.vec("head_new"), st.vec("head_pre")
    mi = int(np.flatnonzero(np.abs(h - hp) == np.abs(h - hp).max())[-1])
    r = recs[-1]
    assert (r["cell_i"], r["cell_j"], r["cell_k"]) == (mi % p.nx + 1, (mi // p.nx) % p.ny + 1,
                                                       mi // (p.nx * p.ny) + 1)
    # FD Jacobian diagonal (or its D-ILU pivot): device-vs-glibc pow ulps
    # divided by the FD step, after k iterations
    np.testing.assert_allclose(r["dmat"], st.vec("dmat")[mi] * L * L, rtol=1e-7)
    np.testing.assert_allclose(r["rhs"], st.vec("rhs")[mi] * L ** 3, rtol=1e-6, atol=1e-20)
    np.testing.assert_allclose(r["head"], h[mi] * L, rtol=0, atol=1e-9)
