"""Look-up-table fluid (FLUID_MODEL = real) on the GPU against the CPU restatement.

The table is the seeded covolume gas held in tests/golden/lut.npz, on which the
restatement itself is pinned bit-exact to the reference's FluidReal (test_oracle_lut.py).
Bars as for the ideal gas: initial state and time step bit-exact, per-call residuals to
1e-13 relative, 60-iteration runs to 1e-11 in the log-residual history, the fields, the
side outputs and the turbomachinery figures.  A table-fluid case advanced in a batch with
ideal-gas cases equals the same case advanced alone, bit for bit.
"""
import numpy as np
import pytest

from cturbobfm_amd import Batch, Solver, SolverError, cases

import oracle as O
from test_oracle_lut import lut_from_fixture

pytestmark = pytest.mark.gpu

RUN_TOL = 1e-11
CALL_TOL = 1e-13


@pytest.fixture(scope="module")
def lut_path(tmp_path_factory):
    import os
    gold = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lut.npz"),
                   allow_pickle=False)
    return cases.write_lut(lut_from_fixture(gold), str(tmp_path_factory.mktemp("lut") / "covolume.lut"))


CASES = {
    "bump2d_jst": lambda: cases.bump2d(33, 17, n_iterations=200),
    "bump2d_roe_muscl_va": lambda: cases.bump2d(33, 17, n_iterations=200, scheme="ROE", muscl=True,
                                                  limiter="van albada", cfl=1.0),
    "rotor3d_jst_hall": lambda: cases.rotor3d(25, 9, 17, n_iterations=200),
    "rotor3d_roe_muscl_liftdrag": lambda: cases.rotor3d(25, 9, 17, n_iterations=200, scheme="ROE", muscl=True,
                                                          limiter="van leer", bfm_model="Lift-Drag", cfl=1.2),
}


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.ndim >= 2 and a.shape[0] == 5:
        return max(np.max(np.abs(a[v] - b[v])) / max(np.max(np.abs(b[v])), 1e-300) for v in range(5))
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def real(name, lut_path):
    return cases.with_real_gas(CASES[name](), lut_path)


@pytest.mark.parametrize("name", list(CASES))
def test_initial_state_and_timestep(name, lut_path):
    case = real(name, lut_path)
    print("case", name, flush=True)
    s = Solver.from_case(case)
    oc = O.OracleCase(case)
    assert np.array_equal(s.state(), oc.state()), "initial state differs"
    U = cases.perturb_state(case, oc.state(), 0.02, 3)
    assert np.array_equal(s.timestep(U), oc.timestep(U)), "time step differs"


@pytest.mark.parametrize("name", list(CASES))
def test_residual_call(name, lut_path):
    case = real(name, lut_path)
    print("case", name, flush=True)
    s = Solver.from_case(case)
    oc = O.OracleCase(case)
    U = cases.perturb_state(case, oc.state(), 0.02, 3)
    oc.set_state(U)
    s.set_state(U)
    Ro = oc.residual(U, it=1)
    Rg = s.residual(U, it=1)
    err = rel_err(Rg, Ro)
    print("residual rel err %.3e, bit-exact fraction %.4f" % (err, np.mean(Rg == Ro)), flush=True)
    assert err <= CALL_TOL


@pytest.mark.parametrize("name", list(CASES))
def test_run_history_and_fields(name, lut_path):
    case = real(name, lut_path)
    print("case", name, flush=True)
    n = 60
    s = Solver.from_case(case)
    oc = O.OracleCase(case)
    lo, _, _ = oc.run(n, threads=8)
    lg = s.advance(n)
    hist_err = np.max(np.abs(lg - lo))
    f_err = rel_err(s.state(), oc.state())
    print("history abs err %.3e, field rel err %.3e" % (hist_err, f_err), flush=True)
    assert hist_err <= RUN_TOL and f_err <= RUN_TOL
    for a, b in zip(s.side_outputs(), oc.side_outputs()):
        if np.max(np.abs(b)) > 0:
            assert np.max(np.abs(a - b)) / np.max(np.abs(b)) <= RUN_TOL


def test_turbo_outputs(lut_path):
    case = real("rotor3d_jst_hall", lut_path)
    s = Solver.from_case(case)
    oc = O.OracleCase(case)
    oc.run(40, threads=8)
    s.advance(40)
    tg, to = s.turbo(), oc.turbo()
    print("turbo gpu", tg[6:], "oracle", to[6:], flush=True)
    assert np.max(np.abs(tg - to) / np.maximum(np.abs(to), 1e-30)) <= RUN_TOL


def test_mixed_precision_refused(lut_path):
    with pytest.raises(SolverError, match="FP64"):
        Solver.from_case(real("bump2d_jst", lut_path), {"GPU_PRECISION": "mixed"})


def test_missing_table_file(tmp_path):
    case = cases.with_real_gas(CASES["bump2d_jst"](), str(tmp_path / "absent.lut"))
    with pytest.raises(SolverError, match="Could not open table file"):
        Solver.from_case(case)


def test_batch_with_ideal_members(lut_path):
    n = 40
    names = ["rotor3d_jst_hall", "bump2d_roe_muscl_va"]
    alone = {}
    for nm in names:
        s = Solver.from_case(real(nm, lut_path))
        s.advance(n)
        alone[nm] = s.state()
        s.close()
    b = Batch()
    members = {nm: Solver.from_case(real(nm, lut_path)) for nm in names}
    ideal = Solver.from_case(CASES["rotor3d_jst_hall"]())
    for s in list(members.values()) + [ideal]:
        b.admit(s)
    b.advance(n)
    for nm, s in members.items():
        assert np.array_equal(s.state(), alone[nm]), "table-fluid %s differs in the mix" % nm
    ref = Solver.from_case(CASES["rotor3d_jst_hall"]())
    ref.advance(n)
    assert np.array_equal(ideal.state(), ref.state()), "ideal member differs in the mix"
    b.close()
