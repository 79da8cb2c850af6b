"""Many domains at once (msg123_amd.ensemble): one context and stream per
domain, host threads overlapping them.  Each domain's results are
bit-identical to its standalone solve."""
import numpy as np
import pytest

from msg123_amd import configs, ensemble

pytestmark = pytest.mark.gpu


def _problems():
    sizes = [(12, 10, 8), (9, 7, 6), (16, 11, 5), (7, 13, 9), (10, 10, 10), (14, 6, 7)]
    return [configs.make_iid(f"e{k}", nx, ny, nz, 10.0, 10.0, 0.5, 500 + k,
                             wt_depth=2.0 + k, recharge_ms=1e-7 * (1 + k), nsteps=4)
            for k, (nx, ny, nz) in enumerate(sizes)]


@pytest.mark.parametrize("mixed", [False, True])
def test_concurrent_equals_sequential(mixed):
    probs = _problems()
    seq = ensemble.run(probs, threads=1, mixed=mixed, keep_heads=True)
    con = ensemble.run(probs, threads=3, mixed=mixed, keep_heads=True)
    print("sequential", round(seq["wall_s"], 3), "s; concurrent", round(con["wall_s"], 3), "s",
          flush=True)
    for a, b in zip(seq["instances"], con["instances"]):
        assert a["name"] == b["name"]
        assert np.array_equal(a["times"], b["times"])
        assert a["newton"] == b["newton"]
        assert np.array_equal(a["heads"], b["heads"])
    assert con["instance_steps"] == sum(r["steps"] for r in con["instances"])


def test_config3_instances_concurrent_equal_standalone():
    """Sixteen instances of config 3 itself (seeds 1000..1015: boxes of 32..128
    x 32..128 x 16..48 cells, their own soil, water table and recharge), two
    hourly steps on the mixed path: solved concurrently, every instance's step
    times, Newton counts and heads equal its standalone solve bit for bit."""
    probs = configs.config3(count=16, nsteps=2)
    seq = ensemble.run(probs, threads=1, mixed=True, keep_heads=True)
    con = ensemble.run(probs, threads=8, mixed=True, keep_heads=True)
    print("cells", sum(p.nx * p.ny * p.nz for p in probs), "sequential",
          round(seq["wall_s"], 3), "s; concurrent", round(con["wall_s"], 3), "s", flush=True)
    for a, b in zip(seq["instances"], con["instances"]):
        assert a["name"] == b["name"]
        assert np.array_equal(a["times"], b["times"])
        assert a["newton"] == b["newton"]
        assert np.array_equal(a["heads"], b["heads"])
