"""msg123_calc_mass_noalloc (the entry point that allocates nothing) against
msg123_calc_mass, through the C ABI: two calls in a row on contexts of
different sizes (the second smaller), then the first again, so the shared
device scratch holds another call's partials each time."""
import sys

import numpy as np
import pytest

from msg123_amd import configs, gpu

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _stepped(p, steps):
    box = gpu.BoxMesh(p)
    ctx = gpu.Context(box)
    ctx.run(nstep_max=steps, keep_heads=False)
    return box, ctx


def test_calc_mass_noalloc_matches_calc_mass():
    _say("test_calc_mass_noalloc_matches_calc_mass")
    big = _stepped(configs.small(40, 33, 17, seed=51, nsteps=3), 2)
    small = _stepped(configs.small(9, 7, 6, seed=11, nsteps=2, dt=7200.0, wt_depth=2.0), 1)
    for name, (box, ctx) in (("big", big), ("small", small), ("big again", big)):
        got = ctx.calc_mass_noalloc()
        want = ctx.calc_mass()
        _say("  ", name, ctx.n, "cells", got)
        # same kernels, same launch shape, same fixed-order fold: identical sums
        for k in want:
            assert got[k] == want[k], (name, k, got[k], want[k])
        assert got["recharge_abs"] > 0.0
    assert gpu.NOALLOC_MAX_CELLS == 65536 * 256
