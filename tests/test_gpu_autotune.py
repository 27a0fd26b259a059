"""The setup's launch-shape autotuner (rows per diagonal-step slice and L21
entries per contribution item, per tree height of every triangular-solve plan,
chosen by timing) must not change any result: PCG histories and solutions are
bit-identical with it on and off (HDD_AUTOTUNE=0)."""
import os

import numpy as np
import pytest

from hierarchicaldd_amd import domain
from tests.gpu_common import same_bits

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["sphere64_sub16", "box48_sub16"])
def test_autotune_bit_identical(case):
    from hierarchicaldd_amd import hdd
    dom, cfg = domain.config0() if case == "sphere64_sub16" else domain.small_box(48, 16, 3, seed=2)
    tol = 1e-6 * float(np.abs(dom.b).max())
    out = {}
    for mode in ("0", "1"):
        os.environ["HDD_AUTOTUNE"] = mode
        try:
            g = hdd.DDSolver(dom, cfg)
        finally:
            os.environ.pop("HDD_AUTOTUNE", None)
        it, hist, _ = g.solve(tol, 300)
        out[mode] = (it, hist, g.get(hdd.CH_X), g.stats()["tuned_launch_shapes"])
        g.close()
    print(f"[{case}] iterations {out['0'][0]} / {out['1'][0]}, tuned launch shapes {out['1'][3]:.0f}")
    assert out["0"][3] == 0
    assert out["0"][0] == out["1"][0] and np.array_equal(out["0"][1], out["1"][1])
    assert same_bits(out["0"][2], out["1"][2])
