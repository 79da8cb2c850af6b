"""Problem definitions of the golden fixtures (shared by make_golden.py and the tests)."""
from msg123_amd import configs

CASES = {
    "small": lambda: configs.small(),                                        # 6x5x4, 3 steps
    "box12": lambda: configs.small(12, 10, 8, seed=5, nsteps=6),             # 960 cells, 6 steps
    "box9": lambda: configs.small(9, 7, 6, seed=11, nsteps=4, dt=7200.0,     # water table 2 m deep
                                  wt_depth=2.0),
}
FULL = ("small", "box12", "box9")
SAMPLED = {"c0": lambda: configs.config0()}
