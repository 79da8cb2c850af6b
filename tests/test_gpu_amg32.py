"""Device AMG preconditioner (MSG123_PRECON_AMG32) against the oracle.

* hierarchy: with the reference's aggregates (complete_orphans = 0) the
  device's patterns equal the reference-order host set-up
  (msg123_amg_setup) element for element, the first prolongator bit for
  bit, and the Galerkin values to FP64 rounding (1e-12 of the largest
  entry: the device sums A P then R (A P), the host in its own passes);
  with the completed aggregation, the same against the oracle's hierarchy;
* V-cycle: the device's FP32 V-cycle equals the oracle's FP64 V-cycle on
  the device's own hierarchy to FP32 accuracy (1e-4 of the largest entry);
* time loop: the preconditioner changes the Krylov path, never the answer:
  heads within 1e-6 m of the reference build at every step.
"""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from msg123_amd import configs, gpu
from oracle import amg as oamg
from oracle import orc
from tests.golden.cases import CASES, FULL

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _system(p, kind="dry"):
    """A system on the box's CSR pattern: the oracle's Jacobian at a dry or a
    ponded (wet) state, or an anisotropic heterogeneous diffusion operator
    (negative diagonal, positive couplings 40x stronger vertically, as the
    layered boxes have) that coarsens through many levels."""
    om = orc.OracleMesh(p)
    s = om.static()
    if kind == "aniso":
        n = om.m.n
        rng = np.random.default_rng(11)
        k = np.exp(rng.normal(0.0, 1.0, n))
        rows = np.repeat(np.arange(n), np.diff(s["offind"]))
        cols = s["offrow"]
        w = np.where(np.abs(cols - rows) >= p.nx * p.ny, 40.0, 1.0)
        lumat = w * np.sqrt(k[rows] * k[cols])
        dmat = -np.bincount(rows, weights=lumat, minlength=n) - 1e-3
        return s["offind"], s["offrow"], rng.standard_normal(n), dmat, lumat
    st = orc.OracleState(om).first_step()
    if kind == "wet":
        st.set_head(s["cell_top"] + 0.02 / s["len_scal"])
    else:
        st.set_head(st.vec("head_new") + np.linspace(-0.05, 0.05, om.m.n))
    rhs, dmat, lumat = st.make_matvec()
    return s["offind"], s["offrow"], rhs, dmat, lumat


def _device_hierarchy(p, dmat, lumat, rhs, **opts):
    box = gpu.BoxMesh(p)
    ctx = gpu.Context(box, precon=gpu.PRECON_AMG32)
    ctx.set_amg_opts(**opts)
    ctx.load_system(rhs, dmat, lumat)
    ctx.precon_dilu()
    return box, ctx, ctx.amg_levels()


def _close(a, b, scale, tol=1e-12):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape
    if a.size:
        assert np.abs(a - b).max() <= tol * scale


@pytest.mark.parametrize("kind", ["dry", "wet", "aniso"])
def test_hierarchy_reference_aggregates(kind):
    """complete_orphans = 0: the device hierarchy is the reference's."""
    _say("test_hierarchy_reference_aggregates", kind)
    p = configs.small(11, 9, 7, seed=19, nsteps=1)
    offind, offrow, rhs, dmat, lumat = _system(p, kind)
    box, ctx, dev = _device_hierarchy(p, dmat, lumat, rhs, complete_orphans=0, coarse_max=1,
                                      project=0)
    ref = gpu.amg_setup(offind, offrow, dmat, lumat, len(dev), 0.05, 0.67)
    _say("  levels", [d["n"] for d in dev], "host", [r["n"] for r in ref])
    assert len(dev) >= (3 if kind == "aniso" else 2)
    for lv in range(1, len(dev)):
        d, r = dev[lv], ref[lv]
        assert d["n"] == r["n"]
        for k in ("offind", "offrow", "pindex", "poffrow"):
            assert np.array_equal(d[k], r[k]), (lv, k)
        scale = max(np.abs(r["dmat"]).max(), np.abs(r["lumat"]).max() if r["nnz"] else 0.0)
        _close(d["dmat"], r["dmat"], scale)
        if r["nnz"]:
            _close(d["lumat"], r["lumat"], scale)
        if lv == 1:
            assert np.array_equal(d["pval"], r["pval"])   # same arithmetic, same order
        else:
            _close(d["pval"], r["pval"], 1.0)
    ctx.close()


@pytest.mark.parametrize("kind,project", [("wet", 0), ("aniso", 0), ("wet", 1), ("dry", 1)])
def test_hierarchy_completed_aggregates_vs_oracle(kind, project):
    """Default aggregation (completed), on the Jacobian or on its M-matrix
    projection."""
    _say("test_hierarchy_completed_aggregates_vs_oracle", kind, project)
    p = configs.small(13, 10, 9, seed=23, nsteps=1)
    offind, offrow, rhs, dmat, lumat = _system(p, kind)
    box, ctx, dev = _device_hierarchy(p, dmat, lumat, rhs, coarse_max=1, project=project)
    if project:
        dmat, lumat = oamg.project(offind, dmat, lumat)
    _, _, ref = oamg.hierarchy(offind, offrow, dmat, lumat, 12, 0.05, 0.67, orphans=True)
    _say("  levels", [d["n"] for d in dev], "oracle", [r["n"] for r in ref])
    assert len(dev) == len(ref) + 1 and len(dev) >= (3 if kind == "aniso" else 2)
    for lv in range(1, len(dev)):
        d, r = dev[lv], ref[lv - 1]
        assert d["n"] == r["n"]
        for k in ("offind", "offrow", "pindex", "poffrow"):
            assert np.array_equal(d[k], r[k]), (lv, k)
        scale = max(np.abs(r["dmat"]).max(), np.abs(r["lumat"]).max() if len(r["lumat"]) else 0)
        _close(d["dmat"], r["dmat"], scale)
        _close(d["lumat"], r["lumat"], scale)
        if lv == 1:
            assert np.array_equal(d["pval"], r["pval"])
        else:
            _close(d["pval"], r["pval"], 1.0)
    ctx.close()


def _mats(p, dmat, lumat, offind, offrow, dev):
    n = len(dmat)
    rows = np.repeat(np.arange(n), np.diff(offind))
    mats = [(sp.csr_matrix((lumat, (rows, offrow)), shape=(n, n)) + sp.diags(dmat)).tocsr()]
    Ps = []
    for lv in range(1, len(dev)):
        d = dev[lv]
        nf = mats[-1].shape[0]
        Ps.append(sp.csr_matrix((d["pval"], d["poffrow"], d["pindex"]), shape=(nf, d["n"])))
        r = np.repeat(np.arange(d["n"]), np.diff(d["offind"]))
        mats.append((sp.csr_matrix((d["lumat"], (r, d["offrow"])), shape=(d["n"], d["n"]))
                     + sp.diags(d["dmat"])).tocsr())
    return mats, Ps


@pytest.mark.parametrize("kind,npre,npost,coarse_max,project", [
    ("aniso", 2, 2, 1, 0), ("aniso", 1, 3, 1, 0), ("aniso", 2, 2, 150, 0), ("aniso", 0, 1, 1, 0),
    ("wet", 2, 2, 1, 0), ("wet", 2, 2, 1, 1), ("dry", 2, 2, 64, 1)])
def test_vcycle_vs_oracle(kind, npre, npost, coarse_max, project):
    """One device V-cycle (FP32) against the oracle's FP64 V-cycle on the
    device's hierarchy; coarse_max = 150 ends on a level solved by 20
    Jacobi sweeps instead of the dense inverse."""
    _say("test_vcycle_vs_oracle", kind, npre, npost, coarse_max, project)
    p = configs.small(12, 11, 10, seed=29, nsteps=1)
    offind, offrow, rhs, dmat, lumat = _system(p, kind)
    box, ctx, dev = _device_hierarchy(p, dmat, lumat, rhs, npre=npre, npost=npost,
                                      coarse_max=coarse_max, project=project)
    if project:
        dmat, lumat = oamg.project(offind, dmat, lumat)
    mats, Ps = _mats(p, dmat, lumat, offind, offrow, dev)
    dense = dev[-1]["n"] <= 80
    _say("  levels", [m.shape[0] for m in mats], "dense coarsest", dense)
    rng = np.random.default_rng(7)
    for trial in range(2):
        r = rng.standard_normal(len(dmat)) * (1.0 + 10.0 * trial)
        z = ctx.solve_ilu(r)
        want = oamg.vcycle(mats, Ps, r, npre=npre, npost=npost, ws=float(np.float32(0.67)),
                           coarse_sweeps=None if dense else 20)
        err = np.abs(z - want).max() / np.abs(want).max()
        _say("  rel err", err)
        assert err <= 1e-4
    ctx.close()


@pytest.mark.parametrize("tag", FULL)
def test_time_loop_vs_reference(tag, golden):
    """AMG-preconditioned BiCGStab with the Newton options the bench runs it
    with (gpu.AMG_NEWTON) on the reference's step times: heads within 1e-6 m
    of the reference build at every step."""
    _say("test_time_loop_vs_reference", tag)
    z = golden(tag)
    box = gpu.BoxMesh(CASES[tag]())
    ctx = gpu.Context(box, precon=gpu.PRECON_AMG32, consistent_jacobian=True)
    ctx.set_newton(**gpu.AMG_NEWTON)
    res = ctx.run(schedule=z["step_times"])
    np.testing.assert_array_equal(res["times"], z["step_times"])
    err = np.abs(res["heads"] - z["step_heads"]).max() * z["len_scal"]
    _say("  max|dh| m", err, "newton", res["newton"].tolist(), "inner", res["inner"].tolist(),
         "set-ups, fallbacks", ctx.amg_stats())
    assert err <= 1e-6
    ctx.close()


def test_refresh_and_options_rebuild_the_hierarchy():
    _say("test_refresh_and_options_rebuild_the_hierarchy")
    p = configs.small(10, 8, 9, seed=31, nsteps=1)
    offind, offrow, rhs, dmat, lumat = _system(p, "aniso")
    box, ctx, dev = _device_hierarchy(p, dmat, lumat, rhs, coarse_max=1)
    r = np.random.default_rng(3).standard_normal(len(dmat))
    z1 = ctx.solve_ilu(r)
    ctx.amg_refresh()
    assert ctx.amg_levels() == [dict(n=0)] or len(ctx.amg_levels()) <= 1
    ctx.precon_dilu()
    z2 = ctx.solve_ilu(r)
    assert np.array_equal(z1, z2)          # deterministic set-up and cycle
    ctx.set_amg_opts(max_levels=2)
    ctx.precon_dilu()
    assert len(ctx.amg_levels()) == 2
    with pytest.raises(RuntimeError):
        ctx.set_amg_opts(npre=-1)
    ctx.close()
