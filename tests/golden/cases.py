"""Problem definitions of the golden fixtures (shared by make_golden.py and the tests)."""
import math

import numpy as np

from msg123_amd import configs

CASES = {
    "small": lambda: configs.small(),                                        # 6x5x4, 3 steps
    "box12": lambda: configs.small(12, 10, 8, seed=5, nsteps=6),             # 960 cells, 6 steps
    "box9": lambda: configs.small(9, 7, 6, seed=11, nsteps=4, dt=7200.0,     # water table 2 m deep
                                  wt_depth=2.0),
    # two inner iterations per linear solve: BiCGStab stops unconverged, so
    # the AMG path's line-search slope (from the last preconditioner input)
    # departs from the D-ILU path's (from rhs)
    "box9_m2": lambda: configs.small(9, 7, 6, seed=36, nsteps=3, dt=7200.0, wt_depth=2.0,
                                     maxinn_iter=2),
}
# recharge twice the mean Ks onto a water table 1 m deep: the top layer
# saturates cell by cell (1, 3, 4, 5 ... of 56 cells over the run), so the
# seepage face (Calc_function.f90:304-391, Make_linearsystem.f90:410-497)
# switches on during the time loop
CASES["seep"] = lambda: configs.small(8, 7, 6, seed=41, nsteps=2, dt=1800.0, wt_depth=1.0,
                                      recharge_ms=1e-5, lnks_mean=math.log(5e-6))


def _rech():
    """Recharge records (a recharge file with a time unit): steps from 600 s
    growing x1.2 up to 3600 s, cut at seven record ends, with a dry spell and
    a pulse of 2e-6 m/s (Time_module.f90:274-330,770-858).  Every step takes
    at most 7 Newton iterations, clear of the heuristic's threshold of 8, so
    a count that differs by one does not change the step sequence.  The last
    record ends at end_time: the reference aborts when it reads past its
    last record (its log message at Time_module.f90:343 writes into an empty
    string)."""
    p = configs.small(10, 8, 6, seed=17, nsteps=1, dt=600.0, wt_depth=5.0)
    p.max_step = 3600.0
    p.end_time = 12 * 3600.0
    p.rech_etime = np.float32([5400.0, 9000.0, 14400.0, 20000.0, 28800.0, 36000.0, 43200.0])
    p.rech_mmh = np.float32([0.36, 1.8, 0.0, 7.2, 0.72, 0.36, 0.18])
    p.recharge_mmh = float(p.rech_mmh[0])
    return p


CASES["rech"] = _rech
FULL = ("small", "box12", "box9", "seep", "rech")
SAMPLED = {"c0": lambda: configs.config0()}

# precon_type = 1 (AMG) runs: set_amg namelist values
_AMG_DEFAULT = dict(amg_nlevel=4, maxvcy_iter=1, max_sweep=1, jac_omega=0.67, amg_theta=0.05)
AMG = {
    "amg12": (lambda: configs.small(12, 10, 8, seed=5, nsteps=6), _AMG_DEFAULT),
    "amg_box9": (lambda: configs.small(9, 7, 6, seed=11, nsteps=4, dt=7200.0, wt_depth=2.0),
                 dict(_AMG_DEFAULT, amg_nlevel=3)),
    "amg_m2": (CASES["box9_m2"], dict(_AMG_DEFAULT, amg_nlevel=3)),
    "amg_aniso": (lambda: configs.small(20, 18, 12, seed=7, nsteps=2, dt=1800.0),
                  dict(_AMG_DEFAULT, amg_nlevel=6, amg_theta=0.01, jac_omega=0.5)),
}

# runs whose head and restart files are kept byte for byte
IO = {"io_box9": CASES["box9"]}

# runs whose mass-balance file (whole domain, accumulated over the run) is kept
MASS = ("box9", "seep", "rech")
