"""The reference's regression cases on the GPU: run to convergence, compare with the
CPU restatement (history and fields <= 1e-11, same converged iteration) and with the
reference's reference.csv (relative L2 < 3e-2, the reference's own bar)."""
import numpy as np
import pytest

from cturbobfm_amd import Solver

import oracle as O
import regression_cases as RC

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", list(RC.CONFIGS))
def test_gpu_regression_case(name):
    print("case", name, flush=True)
    case = RC.load_case(name)
    oc = O.OracleCase(case)
    lo, conv_o, _ = oc.run(40000, threads=8)
    s = Solver.from_case(case)
    s.solve(40000)
    assert s.converged_at == conv_o
    lg = s.history
    assert lg.shape == lo[:conv_o].shape
    hist_err = float(np.max(np.abs(lg - lo[:conv_o])))
    Ug, Uo = s.state(), oc.state()
    f_err = max(float(np.max(np.abs(Ug[v] - Uo[v])) / np.max(np.abs(Uo[v]))) for v in range(5) if np.max(np.abs(Uo[v])) > 0)
    print("converged at %d; history err %.3e; field err %.3e" % (conv_o, hist_err, f_err))
    assert hist_err <= 1e-11 and f_err <= 1e-11
    ref = RC.load_reference_solution(name)
    got = RC.derived_fields(Ug)
    for key in ("Density", "Pressure", "Temperature"):
        assert RC.rel_l2(ref[key], got[key]) < 3e-2
