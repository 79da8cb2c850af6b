"""Generate the golden fixtures in tests/golden/ from the reference build.

Runs oracle/_ref/ref_harness (the reference Fortran solver compiled from its
own sources by oracle/Makefile, plus a dump harness) on decks written by
oracle/deck.py, and stores inputs/outputs as compressed .npz files.  Needs
the reference checkout and amdflang; the fixtures themselves are plain data
and travel with the repository.

    make -C oracle && python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import shutil
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import refdump  # noqa: E402
from oracle.deck import write_deck  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def run(p, tag, full=True, sample=0):
    d = tempfile.mkdtemp(prefix=f"golden_{tag}_")
    try:
        write_deck(p, d)
        refdump.run_reference(d)
        st = refdump.read_static(d)
        a1 = refdump.read_assem1(d, st)
        wet = refdump.read_wet(d, st)
        steps = refdump.read_steps(d, st["n"])
        conv = refdump.read_conv(d)
        out = dict(nx=p.nx, ny=p.ny, nz=p.nz, seed_name=p.name,
                   step_times=steps["times"], step_dts=steps["dts"], step_newton=steps["newton"],
                   step_inner=np.array([sum(it[1] for it in s) for s in conv], np.int32),
                   step_back=np.array([sum(it[2] for it in s) for s in conv], np.int32),
                   len_scal=st["len_scal"], len_scal_inv=st["len_scal_inv"], newper=st["newper"])
        n = st["n"]
        if full:
            for k in ("offind", "offrow", "dir", "hydf", "abyd", "cell_vol", "cell_top", "reta",
                      "retn", "resi", "poro", "ss", "hydf_surf", "abyd_surf", "surf_top",
                      "surf_bott", "rech_area"):
                out["st_" + k] = st[k]
            for k in ("calc_rech", "head_new", "srat_old", "rhs", "dmat", "lumat", "srat_new",
                      "rel_perm", "dh", "bnorm", "rnorm", "in_iter", "conv_flag", "delt"):
                out["a1_" + k] = np.asarray(a1[k])
            for k in ("head", "surf_head", "rhs", "dmat", "lumat"):
                out["wet_" + k] = wet[k]
            out["step_heads"] = steps["heads"]
        else:
            rng = np.random.Generator(np.random.PCG64(12345))
            idx = np.sort(rng.choice(n, size=sample, replace=False))
            out["sample_idx"] = idx.astype(np.int64)
            out["step_heads_sample"] = steps["heads"][:, idx]
            out["step_head_sum"] = steps["heads"].sum(axis=1)
            out["step_head_min"] = steps["heads"].min(axis=1)
            out["step_head_max"] = steps["heads"].max(axis=1)
            out["a1_rhs_sample"] = a1["rhs"][idx]
            out["a1_dmat_sample"] = a1["dmat"][idx]
            out["a1_rhs_norm2"] = np.array(np.dot(a1["rhs"], a1["rhs"]))
        np.savez_compressed(os.path.join(OUT, f"ref_{tag}.npz"), **out)
        print(tag, "steps", len(steps["times"]), "newton", steps["newton"].tolist())
    finally:
        shutil.rmtree(d, ignore_errors=True)


def run_amg(p, tag, amg):
    """precon_type = 1 run: the time loop with the AMG-preconditioned solver
    and the AMG hierarchy of the Jacobian at the final state (levels 1..nlevel:
    CSR system, P, R), with the run's single-thread semantics."""
    d = tempfile.mkdtemp(prefix=f"golden_{tag}_")
    try:
        write_deck(p, d, precon_type=1, amg=amg)
        refdump.run_reference(d)
        st = refdump.read_static(d)
        steps = refdump.read_steps(d, st["n"])
        conv = refdump.read_conv(d)
        h = refdump.read_amg(d, st["n"])
        out = dict(nx=p.nx, ny=p.ny, nz=p.nz, seed_name=p.name, nlevel=h["nlevel"],
                   amg_nlevel=amg["amg_nlevel"], amg_theta=np.float32(amg["amg_theta"]),
                   jac_omega=np.float32(amg["jac_omega"]), amg_head=h["head"],
                   step_times=steps["times"], step_dts=steps["dts"], step_newton=steps["newton"],
                   step_inner=np.array([sum(it[1] for it in s) for s in conv], np.int32),
                   step_back=np.array([sum(it[2] for it in s) for s in conv], np.int32),
                   step_heads=steps["heads"], len_scal=st["len_scal"])
        for lv, L in enumerate(h["levels"], start=1):
            for k, v in L.items():
                out[f"L{lv}_{k}"] = np.asarray(v)
        np.savez_compressed(os.path.join(OUT, f"ref_{tag}.npz"), **out)
        print(tag, "steps", len(steps["times"]), "newton", steps["newton"].tolist(), "levels",
              [L["n"] for L in h["levels"]])
    finally:
        shutil.rmtree(d, ignore_errors=True)


def run_io(p, tag):
    """The reference's own head and restart files (out_list 'head,rest') of a
    run, byte for byte, with the FP64 heads of every accepted step."""
    d = tempfile.mkdtemp(prefix=f"golden_{tag}_")
    try:
        write_deck(p, d)
        refdump.run_reference(d)
        st = refdump.read_static(d)
        steps = refdump.read_steps(d, st["n"])
        files = {}
        for key, prefix in (("head_file", "head_"), ("rest_file", "rest_")):
            name = [f for f in os.listdir(d) if f.startswith(prefix) and f.endswith(".bin")]
            assert len(name) == 1, name
            files[key] = np.fromfile(os.path.join(d, name[0]), np.uint8)
        np.savez_compressed(os.path.join(OUT, f"ref_{tag}.npz"), nx=p.nx, ny=p.ny, nz=p.nz,
                            len_scal=st["len_scal"], step_times=steps["times"],
                            step_heads=steps["heads"], **files)
        print(tag, "head file", files["head_file"].size, "restart file", files["rest_file"].size)
    finally:
        shutil.rmtree(d, ignore_errors=True)


def run_mass(cases, tags):
    """The reference's own mass-balance file (write_out_massf,
    Write_output.f90:497-538: for the whole domain, the terms accumulated
    over the run in m^3, format E13.6) of each case, with its step times."""
    out = {}
    for tag in tags:
        p = cases[tag]()
        d = tempfile.mkdtemp(prefix=f"golden_mass_{tag}_")
        try:
            write_deck(p, d, mass=True)
            refdump.run_reference(d)
            st = refdump.read_static(d)
            steps = refdump.read_steps(d, st["n"])
            name = [f for f in os.listdir(d) if f.startswith("mass_") and f.endswith(".csv")]
            assert len(name) == 1, name
            lines = open(os.path.join(d, name[0])).read().splitlines()
            assert lines[1].split(",")[1:] == ["FLOW", "STORAGE", "RECHARGE", "WELL", "SURFACE",
                                               "RIVER", "LAKE", "SEA", "ALL"], lines[:2]
            rows = lines[2:]
            out[f"{tag}_text"] = np.array(rows)
            out[f"{tag}_values"] = np.array([[float(v) for v in r.split(",")] for r in rows])
            out[f"{tag}_step_times"] = steps["times"]
            out[f"{tag}_len_scal"] = np.array(st["len_scal"])
            print(tag, "mass lines", len(rows), rows[-1])
        finally:
            shutil.rmtree(d, ignore_errors=True)
    np.savez_compressed(os.path.join(OUT, "ref_mass.npz"), **out)


def run_bnd(p, tag, bterms, write_bterms):
    """A run with river, lake, well and sea terms set in the reference's
    module variables (bterms.bin, oracle/ref_harness.f90): static arrays,
    the first assembly and linear solve, the wet state's system, FP64 heads
    of every step and the mass-balance file (its WELL, RIVER, LAKE and SEA
    columns), with the boundary arrays themselves."""
    from oracle import orc
    st0 = orc.OracleMesh(p).static()
    bt = bterms(st0)
    d = tempfile.mkdtemp(prefix=f"golden_{tag}_")
    try:
        write_deck(p, d, mass=True)
        write_bterms(bt, os.path.join(d, "bterms.bin"))
        refdump.run_reference(d)
        st = refdump.read_static(d)
        a1 = refdump.read_assem1(d, st)
        wet = refdump.read_wet(d, st)
        steps = refdump.read_steps(d, st["n"])
        name = [f for f in os.listdir(d) if f.startswith("mass_") and f.endswith(".csv")]
        assert len(name) == 1, name
        rows = open(os.path.join(d, name[0])).read().splitlines()[2:]
        out = dict(nx=p.nx, ny=p.ny, nz=p.nz, step_times=steps["times"], step_dts=steps["dts"],
                   step_newton=steps["newton"], step_heads=steps["heads"],
                   len_scal=st["len_scal"], newper=st["newper"],
                   mass_values=np.array([[float(v) for v in r.split(",")] for r in rows]),
                   mass_text=np.array(rows))
        for k in ("calc_rech", "head_new", "srat_old", "rhs", "dmat", "lumat", "dh", "in_iter",
                  "conv_flag", "delt"):
            out["a1_" + k] = np.asarray(a1[k])
        for k in ("head", "surf_head", "rhs", "dmat", "lumat"):
            out["wet_" + k] = wet[k]
        for k, v in bt.items():
            out["bt_" + k] = v
        np.savez_compressed(os.path.join(OUT, f"ref_{tag}.npz"), **out)
        print(tag, "steps", len(steps["times"]), "newton", steps["newton"].tolist(), "mass",
              rows[-1])
    finally:
        shutil.rmtree(d, ignore_errors=True)


def sample_c1(d, tag="c1", nrand=1024, nwt=2048, nmove=1024):
    """Config 1 (256x256x64, 4.2M cells) from a reference run directory made
    with MSG123_MAX_STEPS=48 (the reference's 48 accepted steps take about 35
    minutes on 8 cores, see run_c1): per-step times, dt, Newton, inner and
    backtracking counts, whole-field head sum / min / max, the same over the
    water-table layer (the cells at pressure head 0, where the residual is
    not differentiable), and per-step heads of 4096 cells: 1024 uniformly
    random, 2048 random cells of the water-table layer and 1024 random cells
    among those that moved more than 1e-6 m over the run."""
    st = refdump.read_static(d)
    n, L = st["n"], st["len_scal"]
    steps = refdump.read_steps(d, n)
    conv = refdump.read_conv(d)
    H = steps["heads"]
    top = st["cell_top"]
    wt = np.flatnonzero(np.abs(H[0] - top) * L < 1e-6)
    moved = np.flatnonzero(np.abs(H[-1] - H[0]) * L > 1e-6)
    rng = np.random.Generator(np.random.PCG64(20261204))
    idx = np.unique(np.concatenate([rng.choice(n, nrand, replace=False),
                                    rng.choice(wt, min(nwt, wt.size), replace=False),
                                    rng.choice(moved, min(nmove, moved.size), replace=False)]))
    out = dict(nx=256, ny=256, nz=64, len_scal=L, newper=st["newper"],
               step_times=steps["times"], step_dts=steps["dts"], step_newton=steps["newton"],
               attempt_inner=np.array([sum(it[1] for it in s) for s in conv], np.int32),
               attempt_newton=np.array([len(s) for s in conv], np.int32),
               sample_idx=idx.astype(np.int64), step_heads_sample=H[:, idx],
               step_head_sum=H.sum(axis=1), step_head_min=H.min(axis=1),
               step_head_max=H.max(axis=1), wt_idx_first=np.int64(wt[0]),
               wt_count=np.int64(wt.size), wt_head_sum=H[:, wt].sum(axis=1),
               wt_head_min=H[:, wt].min(axis=1), wt_head_max=H[:, wt].max(axis=1),
               a1_rhs_norm2=np.array(0.0))
    a1 = refdump.read_assem1(d, st)
    out["a1_rhs_norm2"] = np.array(np.dot(a1["rhs"], a1["rhs"]))
    out["a1_rhs_sample"] = a1["rhs"][idx]
    np.savez_compressed(os.path.join(OUT, f"ref_{tag}.npz"), **out)
    print(tag, "steps", len(steps["times"]), "newton", steps["newton"].tolist(), "sample", idx.size)


def run_c1(d):
    """Run the reference build on config 1 for 48 accepted steps into d (8
    OpenMP threads), then sample it (sample_c1)."""
    import subprocess
    from msg123_amd import configs
    write_deck(configs.config1(), d)
    env = dict(os.environ, OMP_NUM_THREADS="8", OMP_PROC_BIND="close", OMP_PLACES="cores",
               MSG123_MAX_STEPS="48")
    subprocess.run([refdump.HARNESS], cwd=d, check=True, env=env, stdout=subprocess.DEVNULL,
                   stderr=subprocess.DEVNULL)
    sample_c1(d)


if __name__ == "__main__":
    if len(sys.argv) == 3 and sys.argv[1] == "c1":
        # python tests/golden/make_golden.py c1 <run dir>: sample an existing
        # run directory, or run the reference there first when it has none
        d = sys.argv[2]
        if os.path.exists(os.path.join(d, "steps.bin")):
            sample_c1(d)
        else:
            run_c1(d)
        sys.exit(0)
    sys.path.insert(0, OUT)
    from cases import AMG, CASES, IO, MASS, SAMPLED, bterms, write_bterms  # noqa: E402
    only = sys.argv[1:]
    if not only or "bnd" in only:
        run_bnd(CASES["bnd"](), "bnd", bterms, write_bterms)
    if not only or "mass" in only:
        run_mass(CASES, MASS)
    for tag, mk in CASES.items():
        if tag != "bnd" and (not only or tag in only):
            run(mk(), tag)
    for tag, mk in SAMPLED.items():
        if not only or tag in only:
            run(mk(), tag, full=False, sample=400)
    for tag, (mk, amg) in AMG.items():
        if not only or tag in only:
            run_amg(mk(), tag, amg)
    for tag, mk in IO.items():
        if not only or tag in only:
            run_io(mk(), tag)
