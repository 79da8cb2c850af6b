"""Problem definitions of the golden fixtures (shared by make_golden.py and the tests)."""
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
FULL = ("small", "box12", "box9")
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
