"""The body-force part of the solver's source loop (Solver::computeSourceResiduals,
src/solver.cpp:2814-2896) over a whole blade row, against the reference's own
SourceBFMBase::computeTotalSource run in the reference's i, j, k order with one model
object and persistent force / deviation arrays, over three consecutive sweeps
(tests/golden/bfm_sweep_<model>.npz from tests/golden/gen_golden.py sweep).  The
per-node golden vectors of test_oracle_golden.py call the model on isolated nodes; here
any state the model object or the side-output arrays carry from node to node and from
sweep to sweep (a stalled node keeps the viscous force of an earlier sweep) is
exercised.  Agreement is bit-exact."""
import os

import numpy as np
import pytest

from cturbobfm_amd import cases

import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def sweep_case(name):
    if name == "hall":
        return cases.rotor3d(11, 5, 7, bfm_model="Hall")
    case = cases.rotor3d(11, 5, 7, bfm_model="Hall-Thollet", off_design=True, stall=True)
    f = case["fields"]["deviationAngleStall"]
    f[f > 0] = 0.6   # about half of the blade nodes stall, and some switch between sweeps
    return case


@pytest.mark.parametrize("name", ["hall", "thollet_od_stall"])
def test_body_force_sweeps_equal_reference(name):
    g = np.load(os.path.join(GOLD, "bfm_sweep_%s.npz" % name), allow_pickle=False)
    oc = O.OracleCase(sweep_case(name))
    for q in range(len(g["U"])):
        R = oc.source_residual(g["U"][q], which=4)
        assert np.any(R != 0.0)
        assert np.array_equal(R, g["R"][q]), q
        finv, fvisc, dev, _ = oc.side_outputs()
        assert np.array_equal(finv, g["finv"][q]), q
        assert np.array_equal(fvisc, g["fvisc"][q]), q
        assert np.array_equal(dev, g["dev"][q]), q
    if name == "thollet_od_stall":
        # nodes that stall in a later sweep keep an earlier sweep's viscous force
        d = g["dev"]
        switched = (d[1] > 0.6) & (d[0] <= 0.6) & (d[0] != 0.0)
        assert np.any(switched) and np.any((d[0] > 0.0) & (d[0] <= 0.6))
        assert np.array_equal(g["fvisc"][1][switched], g["fvisc"][0][switched])
