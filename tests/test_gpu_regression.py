"""The reference's regression cases on the GPU: run to convergence, compare with the
CPU restatement (same converged iteration, final fields <= 1e-11) and with the
reference's reference.csv (relative L2 < 3e-2, the reference's own bar).

The log-residual history of a run driven 3-4 orders down is ill-conditioned: the
residual is a small difference of large face fluxes, so a 1-ulp difference anywhere
(libm pow in the inlet bisection differs between glibc and the device library in the
last bit for some arguments) grows to ~1e-11 in log10 over 2000 iterations.  The bar for
the history is therefore calibrated on the oracle itself: the same run started from a
state perturbed by 1 ulp gives the oracle's own spread, and the GPU must stay within 4x
of it (and never above 1e-9)."""
import numpy as np
import pytest

from cturbobfm_amd import Solver

import oracle as O
import regression_cases as RC

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", list(RC.CONFIGS))
def test_gpu_regression_case(name, tmp_path):
    print("case", name, flush=True)
    case = RC.load_case(name)
    oc = O.OracleCase(case)
    start = RC.initial_state(name, str(tmp_path / name))
    if start is not None:
        oc.set_state(start)
    U0 = oc.state()
    nmax = RC.max_iterations(name)
    lo, conv_o, _ = oc.run(nmax, threads=8)
    assert (conv_o > 0) == RC.converges(name)
    nrun = conv_o if conv_o else nmax
    # the oracle's own sensitivity: same run from a state 1 ulp away
    op = O.OracleCase(case)
    op.set_state(np.nextafter(U0, np.inf))
    lp, _, _ = op.run(nrun, threads=8)
    spread = float(np.max(np.abs(lp[:nrun] - lo[:nrun])))
    hist_tol = min(max(1e-11, 4.0 * spread), 1e-9)
    s = Solver.from_case(case)
    if start is not None:
        s.set_state(start)
    s.solve(nmax)
    assert s.converged_at == conv_o
    lg = s.history
    assert lg.shape == lo[:nrun].shape
    hist_err = float(np.max(np.abs(lg - lo[:nrun])))
    Ug, Uo = s.state(), oc.state()
    f_err = max(float(np.max(np.abs(Ug[v] - Uo[v])) / np.max(np.abs(Uo[v]))) for v in range(5) if np.max(np.abs(Uo[v])) > 0)
    print("converged at %d (of %d); history err %.3e (oracle 1-ulp spread %.3e); field err %.3e"
          % (conv_o, nmax, hist_err, spread, f_err))
    assert hist_err <= hist_tol and f_err <= 1e-11
    ref = RC.load_reference_solution(name)
    got = RC.derived_fields(Ug)
    for key in ("Density", "Pressure", "Temperature"):
        assert RC.rel_l2(ref[key], got[key]) < 3e-2
