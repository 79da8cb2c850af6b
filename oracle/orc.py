"""ctypes wrapper of the C oracle (oracle/orc.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, always as the checker, never as the product.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liborc.so")

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)


class Mesh(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("n", C.c_int),
                ("ns", C.c_int), ("nconn", C.c_int),
                ("len_scal", C.c_double), ("len_scal_inv", C.c_double),
                ("newper", C.c_double), ("newper_inv", C.c_double),
                ("offind", _ip), ("offrow", _ip), ("dir", _ip),
                ("hydf", _dp), ("abyd", _dp), ("cell_vol", _dp), ("cell_top", _dp),
                ("reta", _fp), ("retn", _fp), ("resi", _fp), ("poro", _fp), ("ss", _fp),
                ("hydf_surf", _dp), ("abyd_surf", _dp), ("surf_top", _dp),
                ("surf_bott", _dp), ("rech_area", _dp), ("read_rech", _fp), ("h0", _dp),
                ("bt", C.c_void_p)]


class BTerms(C.Structure):
    _fields_ = [("nrive", C.c_int), ("rive_cell", _ip), ("rive_head", _dp), ("rive_bott", _dp),
                ("rive_abyd", _dp), ("nlake", C.c_int), ("lake_cell", _ip), ("lake_head", _dp),
                ("lake_bott", _dp), ("lake_abyd", _dp), ("nwell", C.c_int), ("well_conn", _ip),
                ("calc_well", _dp), ("nseal", C.c_int), ("seal_cell", _ip), ("seal_sea", _ip),
                ("seal_hydf", _dp), ("seal_abyd", _dp), ("read_seal", _fp)]


class Settings(C.Structure):
    _fields_ = [("maxout_iter", C.c_int), ("maxinn_iter", C.c_int), ("criteria", C.c_double),
                ("init_step", C.c_float), ("max_step", C.c_float), ("end_time", C.c_float),
                ("precon_type", C.c_int), ("rech_etime", C.POINTER(C.c_float)),
                ("rech_mmh", C.POINTER(C.c_float)), ("n_rech", C.c_int)]


class State(C.Structure):
    _fields_ = [("head_new", _dp), ("head_old", _dp), ("head_pre", _dp), ("head_change", _dp),
                ("srat_new", _dp), ("srat_old", _dp), ("rel_perm", _dp),
                ("surf_head", _dp), ("surf_old", _dp), ("surf_rati", _dp),
                ("calc_rech", _dp), ("head_old2", _dp),
                ("rhs", _dp), ("dmat", _dp), ("lumat", _dp),
                ("delt", C.c_double), ("delt_inv", C.c_double), ("delt_old1", C.c_double),
                ("delt_old2", C.c_double), ("current_t", C.c_float), ("next_time", C.c_float),
                ("conv_flag", C.c_int), ("out_iter", C.c_int), ("timestep_num", C.c_int),
                ("in_iter_total", C.c_int), ("back_total", C.c_int),
                ("last_in_iter", C.c_int), ("last_back_iter", C.c_int),
                ("rk", C.c_int), ("rflag", C.c_int), ("etime", C.c_float)]


def build(force: bool = False) -> str:
    src = [os.path.join(HERE, "orc.c")]
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < max(
            os.path.getmtime(s) for s in src + [os.path.join(HERE, "orc.h")]):
        subprocess.check_call(["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off",
                               "-fno-fast-math", "-o", LIB] + src + ["-lm"])
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()  # also when the library is older than its sources
        L = C.CDLL(LIB)
        L.orc_mesh_box.argtypes = [C.POINTER(Mesh), C.c_int, C.c_int, C.c_int, C.c_float,
                                   C.c_float, C.c_float, _fp, _fp, _fp, _fp, _fp, _fp,
                                   C.c_float, C.c_float]
        L.orc_mesh_free.argtypes = [C.POINTER(Mesh)]
        L.orc_state_init.argtypes = [C.POINTER(State), C.POINTER(Mesh)]
        L.orc_state_free.argtypes = [C.POINTER(State)]
        L.orc_srat_rperm.argtypes = [C.POINTER(Mesh), C.c_double, _dp, _dp, _dp, _dp]
        L.orc_calc_func.argtypes = [C.POINTER(Mesh), C.POINTER(State), _dp, _dp]
        L.orc_make_matvec.argtypes = [C.POINTER(Mesh), C.POINTER(State)]
        L.orc_calc_vecjacf.argtypes = [C.POINTER(Mesh), C.POINTER(State), _dp, _dp, _dp]
        L.orc_precon_dilu.argtypes = [C.POINTER(Mesh), _dp, _dp, _dp]
        L.orc_solve_ilu.argtypes = [C.POINTER(Mesh), _dp, _dp, _dp, _dp]
        L.orc_matvec.argtypes = [C.POINTER(Mesh), _dp, _dp, _dp, _dp]
        L.orc_dot.argtypes = [C.c_int, _dp, _dp]
        L.orc_dot.restype = C.c_double
        L.orc_bicgstab.argtypes = [C.POINTER(Mesh), _dp, _dp, _dp, C.c_double, C.c_double,
                                   C.c_int, _dp, _dp, _ip]
        L.orc_bicgstab.restype = C.c_int
        L.orc_calc_numsol.argtypes = [C.POINTER(Mesh), C.POINTER(State), C.POINTER(Settings)]
        L.orc_update_tstep.argtypes = [C.POINTER(Mesh), C.POINTER(State), C.POINTER(Settings)]
        L.orc_run.argtypes = [C.POINTER(Mesh), C.POINTER(Settings), C.c_int, _dp, _dp, _dp,
                              _ip, _ip]
        L.orc_run.restype = C.c_int
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def arr(ptr, n, dtype=np.float64):
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class OracleMesh:
    """Owns an orc_mesh built for a configs.Problem."""

    def __init__(self, p):
        self.p = p
        self.m = Mesh()
        self._keep = [np.ascontiguousarray(x, dtype=np.float32) for x in
                      (p.alpha, p.vg_n, p.theta_r, p.theta_s, p.ks, p.ss)]
        rc = lib().orc_mesh_box(C.byref(self.m), p.nx, p.ny, p.nz, p.dx, p.dy, p.dz,
                                *[_p(x, C.c_float) for x in self._keep],
                                p.wt_depth, p.recharge_mmh)
        assert rc == 0

    def __del__(self):
        try:
            lib().orc_mesh_free(C.byref(self.m))
        except Exception:
            pass

    def set_bterms(self, bt):
        """River, lake, well and sea terms (tests/golden/cases.py bterms()
        layout: 0-based cells); None removes them."""
        if bt is None:
            self.m.bt = None
            self._bt = None
            return
        i = lambda k: np.ascontiguousarray(bt[k], np.int32)  # noqa: E731
        f = lambda k: np.ascontiguousarray(bt[k], np.float64)  # noqa: E731
        keep = dict(rive_cell=i("rive_cell"), rive_head=f("rive_head"), rive_bott=f("rive_bott"),
                    rive_abyd=f("rive_abyd"), lake_cell=i("lake_cell"), lake_head=f("lake_head"),
                    lake_bott=f("lake_bott"), lake_abyd=f("lake_abyd"), well_conn=i("well_conn"),
                    calc_well=f("calc_well"), seal_cell=i("seal_cell"), seal_sea=i("seal_sea"),
                    seal_hydf=f("seal_hydf"), seal_abyd=f("seal_abyd"),
                    read_seal=np.ascontiguousarray(bt["read_seal"], np.float32))
        t = BTerms()
        t.nrive, t.nlake = len(keep["rive_cell"]), len(keep["lake_cell"])
        t.nwell, t.nseal = len(keep["well_conn"]), len(keep["seal_cell"])
        for k, v in keep.items():
            setattr(t, k, _p(v, {np.int32: C.c_int, np.float64: C.c_double,
                                 np.float32: C.c_float}[v.dtype.type]))
        self._bt = (t, keep)
        self.m.bt = C.cast(C.pointer(t), C.c_void_p)

    def static(self) -> dict:
        m = self.m
        n, ns, nc = m.n, m.ns, m.nconn
        return dict(
            n=n, ns=ns, nconn=nc, len_scal=m.len_scal, len_scal_inv=m.len_scal_inv,
            newper=m.newper,
            offind=arr(m.offind, n + 1, np.int32), offrow=arr(m.offrow, nc, np.int32),
            dir=arr(m.dir, nc, np.int32), hydf=arr(m.hydf, nc), abyd=arr(m.abyd, nc),
            cell_vol=arr(m.cell_vol, n), cell_top=arr(m.cell_top, n),
            reta=arr(m.reta, n, np.float32), retn=arr(m.retn, n, np.float32),
            resi=arr(m.resi, n, np.float32), poro=arr(m.poro, n, np.float32),
            ss=arr(m.ss, n, np.float32),
            hydf_surf=arr(m.hydf_surf, ns), abyd_surf=arr(m.abyd_surf, ns),
            surf_top=arr(m.surf_top, ns), surf_bott=arr(m.surf_bott, ns),
            rech_area=arr(m.rech_area, ns), read_rech=arr(m.read_rech, ns, np.float32),
            h0=arr(m.h0, n))

    def settings(self, precon_type=0) -> Settings:
        p = self.p
        et = getattr(p, "rech_etime", None)
        if et is None:
            return Settings(p.maxout_iter, p.maxinn_iter, p.criteria, p.init_step, p.max_step,
                            p.end_time, precon_type, None, None, 0)
        # kept alive on the mesh object for as long as the settings are used
        self._rech = (np.ascontiguousarray(et, np.float32),
                      np.ascontiguousarray(p.rech_mmh, np.float32))
        return Settings(p.maxout_iter, p.maxinn_iter, p.criteria, p.init_step, p.max_step,
                        p.end_time, precon_type, _p(self._rech[0], C.c_float),
                        _p(self._rech[1], C.c_float), len(self._rech[0]))

    def run(self, nstep_max=100000, precon_type=0):
        n = self.m.n
        cap = nstep_max
        heads = np.zeros(cap * n)
        times = np.zeros(cap)
        dts = np.zeros(cap)
        newton = np.zeros(cap, np.int32)
        inner = np.zeros(cap, np.int32)
        cfg = self.settings(precon_type)
        k = lib().orc_run(C.byref(self.m), C.byref(cfg), cap, _p(heads, C.c_double),
                          _p(times, C.c_double), _p(dts, C.c_double), _p(newton, C.c_int),
                          _p(inner, C.c_int))
        k = min(k, cap) if k >= 0 else k
        return dict(nstep=k, heads=heads[:max(k, 0) * n].reshape(-1, n), times=times[:k],
                    dts=dts[:k], newton=newton[:k], inner=inner[:k])


class OracleState:
    """An orc_state positioned at the first Newton iteration of step 1."""

    def __init__(self, om: OracleMesh):
        self.om = om
        self.s = State()
        lib().orc_state_init(C.byref(self.s), C.byref(om.m))
        self.cfg = om.settings()

    def __del__(self):
        try:
            lib().orc_state_free(C.byref(self.s))
        except Exception:
            pass

    def first_step(self):
        lib().orc_update_tstep(C.byref(self.om.m), C.byref(self.s), C.byref(self.cfg))
        return self

    def set_head(self, h):
        h = np.ascontiguousarray(h, dtype=np.float64)
        np.ctypeslib.as_array(self.s.head_new, shape=(self.om.m.n,))[:] = h

    def calc_vecjacf(self, x, v):
        x = np.ascontiguousarray(x, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        out = np.zeros(self.om.m.n)
        lib().orc_calc_vecjacf(C.byref(self.om.m), C.byref(self.s), x.ctypes.data_as(_dp),
                               v.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
        return out

    def make_matvec(self):
        lib().orc_make_matvec(C.byref(self.om.m), C.byref(self.s))
        m = self.om.m
        return (arr(self.s.rhs, m.n), arr(self.s.dmat, m.n), arr(self.s.lumat, m.nconn))

    def calc_func(self, h):
        h = np.ascontiguousarray(h, dtype=np.float64)
        f = np.zeros(self.om.m.n)
        lib().orc_calc_func(C.byref(self.om.m), C.byref(self.s), _p(h, C.c_double),
                            _p(f, C.c_double))
        return f

    def vec(self, name, n=None):
        return arr(getattr(self.s, name), n if n is not None else self.om.m.n)


def bicgstab(om: OracleMesh, dmat, lumat, rhs, bnorm, errtol, maxinn, x0=None):
    n = om.m.n
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    rn = C.c_double(0.0)
    conv = C.c_int(0)
    dmat = np.ascontiguousarray(dmat, np.float64)
    lumat = np.ascontiguousarray(lumat, np.float64)
    rhs = np.ascontiguousarray(rhs, np.float64)
    it = lib().orc_bicgstab(C.byref(om.m), _p(dmat, C.c_double), _p(lumat, C.c_double),
                            _p(rhs, C.c_double), bnorm, errtol, maxinn, _p(x, C.c_double),
                            C.byref(rn), C.byref(conv))
    return x, it, rn.value, conv.value


def precon_dilu(om, dmat, lumat):
    d = np.zeros(om.m.n)
    dmat = np.ascontiguousarray(dmat, np.float64)
    lumat = np.ascontiguousarray(lumat, np.float64)
    lib().orc_precon_dilu(C.byref(om.m), _p(dmat, C.c_double), _p(lumat, C.c_double),
                          _p(d, C.c_double))
    return d


def solve_ilu(om, r, d, lumat):
    z = np.zeros(om.m.n)
    r = np.ascontiguousarray(r, np.float64)
    d = np.ascontiguousarray(d, np.float64)
    lumat = np.ascontiguousarray(lumat, np.float64)
    lib().orc_solve_ilu(C.byref(om.m), _p(r, C.c_double), _p(d, C.c_double),
                        _p(lumat, C.c_double), _p(z, C.c_double))
    return z


def matvec(om, dmat, lumat, x):
    y = np.zeros(om.m.n)
    args = [np.ascontiguousarray(a, np.float64) for a in (dmat, lumat, x)]
    lib().orc_matvec(C.byref(om.m), *[_p(a, C.c_double) for a in args], _p(y, C.c_double))
    return y
