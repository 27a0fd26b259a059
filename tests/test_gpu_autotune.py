"""The setup's launch-shape autotuner (rows per diagonal-step slice and L21
entries per contribution item, per tree height of every triangular-solve plan,
chosen by timing; a workgroup or a wavefront per run on fused heights) must not
change any result: PCG histories and solutions are bit-identical with it on and
off (HDD_AUTOTUNE=0), and with the wavefront form forced on and off
(HDD_FUSED_WAVE=1 / 0) on every fused height."""
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


@pytest.mark.parametrize("case", ["sphere96_sub32", "box128_sub32"])
def test_fused_wave_form_bit_identical(case):
    """Problems with fused heights (>= 1024 supernodes at a tree height): 1 at 96^3, 5 at 128^3."""
    from hierarchicaldd_amd import hdd
    dom, cfg = (domain.sphere_case(96, 32, 4, delta=(32, 32, 32)) if case == "sphere96_sub32"
                else domain.small_box(128, 32, 4, seed=3))
    tol = 1e-6 * float(np.abs(dom.b).max())
    out = {}
    for mode in ("0", "1"):
        os.environ["HDD_FUSED_WAVE"] = mode
        try:
            g = hdd.DDSolver(dom, cfg)
        finally:
            os.environ.pop("HDD_FUSED_WAVE", None)
        it, hist, _ = g.solve(tol, 300)
        st = g.stats()
        out[mode] = (it, hist, g.get(hdd.CH_X), st["fused_heights"], st["wave_heights"])
        g.close()
    print(f"[{case}] iterations {out['0'][0]} / {out['1'][0]}, fused heights {out['1'][3]:.0f}, wavefront form {out['0'][4]:.0f} / {out['1'][4]:.0f}")
    assert out["1"][3] > 0 and out["1"][4] == out["1"][3] and out["0"][4] == 0
    assert out["0"][0] == out["1"][0] and np.array_equal(out["0"][1], out["1"][1])
    assert same_bits(out["0"][2], out["1"][2])
