"""The reference program with its modules swapped for the product's Fortran
shims (msg123_amd/fortran), against the unmodified reference build, on the
same decks.  The binaries are built by `make -C oracle all dropin` from the
reference sources; the tests skip where they are absent.

* CPU: the AMG set-up shim (make_amg_matrix -> msg123_amg_setup) inside the
  reference's own solver: every output file byte for byte.
* GPU: the Newton shim (nonlinear_solution -> msg123_calc_numsol) inside the
  reference's main program: same step times and Newton counts, heads in the
  float32 head file within 1e-6 m, convergence log with the same lines.
* GPU: the reference's own Newton loop (its Nonlinear_solution.f90) over the
  residual, Jacobian and linear-solver shims (calc_function -> msg123_calc_func /
  msg123_calc_vecjacf, make_linearsystem -> msg123_make_matvec,
  linear_solution -> msg123_solve_linalg): the same bars.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from oracle.deck import write_deck
from tests.golden.cases import AMG, CASES

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")


def _exe(name):
    path = os.path.join(REF, name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not built (make -C oracle all dropin)")
    return path


def _run(exe, p, d, precon_type=0, amg=None):
    shutil.rmtree(d, ignore_errors=True)
    os.makedirs(d)
    write_deck(p, d, precon_type=precon_type, amg=amg)
    subprocess.run([exe], cwd=d, check=True, timeout=600, stdout=subprocess.DEVNULL,
                   stderr=subprocess.DEVNULL, env=dict(os.environ, OMP_NUM_THREADS="1"))
    return d


def _heads(d, n):
    """Head output: per record a float32 time header then n float32 heads
    (Write_module.f90:146-225, stream records)."""
    b = np.fromfile(os.path.join(d, "head_tran_run.bin"), dtype="<f4")
    return b.reshape(-1, n + 1)


_LINE = re.compile(r"^\s+(\d+)\s+(\d+)\s+(\d+)\s+(\S+)\s+(\(\d+,\d+,\d+\))")


def _conv(d):
    steps, cur = [], None
    for line in open(os.path.join(d, "conv_tran_run.txt")):
        if "CURRENT TIME" in line:
            cur = []
            steps.append((line.split(":")[1].split()[0], cur))
        m = _LINE.match(line)
        if m and cur is not None:
            cur.append(m.groups())
        if line.startswith(" Stop") or line.startswith(" Didn't"):
            cur.append(("end", line.strip()))
    return steps


@pytest.mark.parametrize("tag", ["amg_aniso", "amg_m2"])
def test_amg_shim_in_reference_solver(tag, tmp_path):
    """precon_type = 1 runs: the reference solver with the AMG set-up shim
    writes the same hierarchy dump, steps, convergence log and head file."""
    ref, shim = _exe("ref_harness"), _exe("ref_harness_amgshim")
    mk, amg = AMG[tag]
    a = _run(ref, mk(), str(tmp_path / "ref"), 1, amg)
    b = _run(shim, mk(), str(tmp_path / "shim"), 1, amg)
    for f in ("amg.bin", "steps.bin", "conv_tran_run.txt", "head_tran_run.bin"):
        assert open(os.path.join(a, f), "rb").read() == open(os.path.join(b, f), "rb").read(), f


@pytest.mark.gpu
@pytest.mark.parametrize("tag,precon_type", [("box12", 0), ("box9", 0), ("amg12", 1)])
def test_gpu_newton_shim_in_reference_program(tag, precon_type, tmp_path):
    """The reference's main program (Msg.f90) with calc_numsol on the GPU
    against the reference build on the same deck."""
    ref, dropin = _exe("msg123"), _exe("msg123_dropin")
    if precon_type == 1:
        mk, amg = AMG[tag]
        p = mk()
    else:
        p, amg = CASES[tag](), None
    print("dropin", tag, precon_type, flush=True)
    a = _run(ref, p, str(tmp_path / "ref"), precon_type, amg)
    b = _run(dropin, p, str(tmp_path / "gpu"), precon_type, amg)
    n = p.nx * p.ny * p.nz
    ha, hb = _heads(a, n), _heads(b, n)
    assert ha.shape == hb.shape
    assert np.array_equal(ha[:, 0], hb[:, 0])                   # output times
    assert np.abs(ha[:, 1:] - hb[:, 1:]).max() <= 1e-6
    ca, cb = _conv(a), _conv(b)
    assert [t for t, _ in ca] == [t for t, _ in cb]             # every step attempt
    for (_, la), (_, lb) in zip(ca, cb):
        assert abs(len(la) - len(lb)) <= 1                      # Newton count +-1
        assert [r[0] for r in la if r[0] == "end"] == [r[0] for r in lb if r[0] == "end"]


@pytest.mark.gpu
@pytest.mark.parametrize("tag", ["box12", "box9", "seep"])
def test_gpu_residual_jacobian_solver_shims_in_reference_program(tag, tmp_path):
    """Msg.f90 and the reference's calc_numsol with calc_func, calc_vecjacf,
    make_matvec and solve_linalg on the GPU, against the reference build."""
    ref, parts = _exe("msg123"), _exe("msg123_dropin_parts")
    p = CASES[tag]()
    print("dropin parts", tag, flush=True)
    a = _run(ref, p, str(tmp_path / "ref"))
    b = _run(parts, p, str(tmp_path / "gpu"))
    n = p.nx * p.ny * p.nz
    ha, hb = _heads(a, n), _heads(b, n)
    assert ha.shape == hb.shape
    assert np.array_equal(ha[:, 0], hb[:, 0])
    assert np.abs(ha[:, 1:] - hb[:, 1:]).max() <= 1e-6
    ca, cb = _conv(a), _conv(b)
    assert [t for t, _ in ca] == [t for t, _ in cb]
    for (_, la), (_, lb) in zip(ca, cb):
        assert abs(len(la) - len(lb)) <= 1
