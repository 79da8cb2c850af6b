"""Problem definitions of the golden fixtures (shared by make_golden.py and the tests)."""
import math

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
FULL = ("small", "box12", "box9", "seep")
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
