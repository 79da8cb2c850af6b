"""Readers for the reference harness dumps (oracle/ref_harness.f90) and a
runner that produces them.  TEST INFRASTRUCTURE ONLY."""
from __future__ import annotations

import os
import re
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HARNESS = os.path.join(HERE, "_ref", "ref_harness")


class _R:
    def __init__(self, path):
        self.b = open(path, "rb").read()
        self.o = 0

    def take(self, dtype, count):
        a = np.frombuffer(self.b, dtype=dtype, count=count, offset=self.o)
        self.o += a.nbytes
        return a.copy()


def read_static(d):
    r = _R(os.path.join(d, "static.bin"))
    n, ns, nconn, nrech = r.take("<i4", 4)
    ls, lsi, newper = r.take("<f8", 3)
    out = dict(n=int(n), ns=int(ns), nconn=int(nconn), nrech=int(nrech), len_scal=ls,
               len_scal_inv=lsi, newper=newper)
    out["offind"] = r.take("<i4", n + 1)
    out["offrow"] = r.take("<i4", nconn) - 1  # 0-based
    out["dir"] = r.take("<i4", nconn)
    out["hydf"] = r.take("<f8", nconn)
    out["abyd"] = r.take("<f8", nconn)
    out["cell_vol"] = r.take("<f8", n)
    out["cell_top"] = r.take("<f8", n)
    for k in ("reta", "retn", "resi"):
        out[k] = r.take("<f4", n)
    for k in ("poro", "ss"):
        out[k] = r.take("<f4", n)
    for k in ("hydf_surf", "abyd_surf", "surf_top", "surf_bott"):
        out[k] = r.take("<f8", ns)
    out["rech_area"] = r.take("<f8", ns)
    return out


def read_assem1(d, st):
    n, nc, nr = st["n"], st["nconn"], st["nrech"]
    r = _R(os.path.join(d, "assem1.bin"))
    out = {}
    out["delt"] = r.take("<f8", 1)[0]
    out["calc_rech"] = r.take("<f8", nr)
    out["rech2cals"] = r.take("<i4", nr) - 1
    for k in ("head_new", "head_old", "srat_old"):
        out[k] = r.take("<f8", n)
    out["rhs"] = r.take("<f8", n)
    out["dmat"] = r.take("<f8", n)
    out["lumat"] = r.take("<f8", nc)
    out["srat_new"] = r.take("<f8", n)
    out["rel_perm"] = r.take("<f8", n)
    out["in_iter"], out["conv_flag"] = (int(x) for x in r.take("<i4", 2))
    out["bnorm"], out["rnorm"] = r.take("<f8", 2)
    out["dh"] = r.take("<f8", n)
    return out


def read_wet(d, st):
    n, ns, nc = st["n"], st["ns"], st["nconn"]
    r = _R(os.path.join(d, "wet.bin"))
    out = {"head": r.take("<f8", n), "surf_head": r.take("<f8", ns)}
    out["rhs"] = r.take("<f8", n)
    out["dmat"] = r.take("<f8", n)
    out["lumat"] = r.take("<f8", nc)
    return out


def read_steps(d, n):
    b = open(os.path.join(d, "steps.bin"), "rb").read()
    rec = 8 + 8 + 4 + 8 * n
    k = len(b) // rec
    times, dts, newton, heads = [], [], [], []
    for i in range(k):
        o = i * rec
        times.append(np.frombuffer(b, "<f8", 1, o)[0])
        dts.append(np.frombuffer(b, "<f8", 1, o + 8)[0])
        newton.append(int(np.frombuffer(b, "<i4", 1, o + 16)[0]))
        heads.append(np.frombuffer(b, "<f8", n, o + 20).copy())
    return dict(times=np.array(times), dts=np.array(dts), newton=np.array(newton),
                heads=np.array(heads))


_ITER = re.compile(r"^\s+(\d+)\s+(\d+)\s+(\d+)\s+\S+")


def read_conv(d, name="run"):
    """Per step: list of (outer, inner, back) from the reference conv log
    (format Nonlinear_solution.f90:59-63,250-252)."""
    steps, cur = [], None
    for line in open(os.path.join(d, f"conv_tran_{name}.txt")):
        if "CURRENT TIME" in line:
            if cur is not None:
                steps.append(cur)
            cur = []
        m = _ITER.match(line)
        if m and cur is not None:
            cur.append(tuple(int(x) for x in m.groups()))
    if cur is not None:
        steps.append(cur)
    return steps


def run_reference(deck_dir, timeout=3600, threads=1):
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    subprocess.run([HARNESS], cwd=deck_dir, check=True, timeout=timeout, env=env,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)


def read_amg(d, n):
    """AMG hierarchy dumped by the harness for precon_type = 1 runs: per level
    the CSR system (0-based columns) and, from level 2, P and R in CSR."""
    r = _R(os.path.join(d, "amg.bin"))
    nlevel = int(r.take("<i4", 1)[0])
    out = {"nlevel": nlevel, "head": r.take("<f8", n), "levels": []}
    for lv in range(1, nlevel + 1):
        nc, nnz = (int(x) for x in r.take("<i4", 2))
        L = {"n": nc, "nnz": nnz, "offind": r.take("<i4", nc + 1),
             "offrow": r.take("<i4", nnz) - 1, "dmat": r.take("<f8", nc),
             "lumat": r.take("<f8", nnz)}
        if lv >= 2:
            nf = out["levels"][-1]["n"]
            npv = int(r.take("<i4", 1)[0])
            L["pindex"] = r.take("<i4", nf + 1)
            L["poffrow"] = r.take("<i4", npv) - 1
            L["pval"] = r.take("<f8", npv)
            nrv = int(r.take("<i4", 1)[0])
            L["rindex"] = r.take("<i4", nc + 1)
            L["roffrow"] = r.take("<i4", nrv) - 1
            L["rval"] = r.take("<f8", nrv)
        out["levels"].append(L)
    return out
