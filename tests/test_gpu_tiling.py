"""The stage kernel's tilings (GPU_STAGE_TILING: tile shape and k pieces) against one
another.  The residual of a node does not depend on how the grid is cut into tiles and
k pieces, so every tiling gives the same bits: 32x8, 16x16 and 28x9 run the static-schedule
3-D JST kernel, 64x4 (halo beyond one position per thread) the earlier loop, and the
pieces change where each column's queue is warmed up.  The same holds with FP32 flux
(GPU_PRECISION = mixed), where both kernels extrapolate the i/j ghost positions in float
from the state rounded to float.  FP64 is also held to the oracle."""
import numpy as np
import pytest

from cturbobfm_amd import Solver, cases

import oracle as O

pytestmark = pytest.mark.gpu

TILINGS = ["28x9/z1", "28x9/z5", "32x8/z3", "16x16/z2", "64x4/z4"]
CASES = {
    "rotor_jst_hall_29x10x9": lambda: cases.rotor3d(29, 10, 9, n_iterations=100),
    "rotor_jst_thollet_49x17x13": lambda: cases.rotor3d(49, 17, 13, n_iterations=100, bfm_model="Hall-Thollet"),
}


def run(case, tiling, prec, U):
    ov = {"GPU_STAGE_TILING": tiling, "GPU_PRECISION": prec}
    s = Solver.from_case(case, ov)
    s.set_state(U)
    R = s.residual(U)
    s.advance(20)
    V = s.state()
    t = s.tuning()
    s.close()
    assert t.startswith("fixed -> " + tiling.split("/")[0]), t
    return R, V


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_tilings_agree(name, prec):
    case = CASES[name]()
    oc = O.OracleCase(case)
    U = cases.perturb_state(case, oc.state(), 0.02, 5)
    out = {t: run(case, t, prec, U) for t in TILINGS}
    print("case", name, prec, flush=True)
    R0, V0 = out[TILINGS[0]]
    for t in TILINGS[1:]:
        R, V = out[t]
        same_r, same_v = np.array_equal(R, R0), np.array_equal(V, V0)
        print("%-9s residual %s, 20 iterations %s" % (t, "identical" if same_r else "differs",
                                                       "identical" if same_v else "differ"), flush=True)
        assert same_r and same_v, t
    if prec == "fp64":
        oc.set_state(U)
        Ro = oc.residual(U, it=1)
        err = np.max(np.abs(R0 - Ro)) / np.max(np.abs(Ro))
        print("against the oracle: rel err %.3e" % err, flush=True)
        assert err <= 1e-13


def test_tiling_key_refuses_unknown_shapes():
    with pytest.raises(Exception, match="GPU_STAGE_TILING"):
        Solver.from_case(cases.rotor3d(9, 4, 5, n_iterations=10), {"GPU_STAGE_TILING": "48x5"})
