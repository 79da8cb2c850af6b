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


if __name__ == "__main__":
    sys.path.insert(0, OUT)
    from cases import AMG, CASES, IO, SAMPLED  # noqa: E402
    only = sys.argv[1:]
    for tag, mk in CASES.items():
        if not only or tag in only:
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
