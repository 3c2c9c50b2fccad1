"""Viscous cases in batches, and the settings the GPU path refuses.

- A viscous case with a moving no-slip wall (its first stage reads the wall-momentum
  copy) computes the same bits alone and batched with an inviscid case and a viscous
  case at rest.
- TURBULENCE_ACTIVE is refused at set-up with a message; so is the volume file of a
  viscous case.
"""
import numpy as np
import pytest

from cturbobfm_amd import Batch, Solver, SolverError, cases

pytestmark = pytest.mark.gpu


def moving():
    return cases.viscous(cases.bump2d(33, 17, n_iterations=200, scheme="ROE", cfl=0.8),
                         walls={"LOWER_WALL": (40.0, 0.0, 0.0), "UPPER_WALL": (0.0, 0.0, 0.0)})


def duct():
    case = cases.rotor3d(25, 9, 17, n_iterations=200)
    case["config"].update({"BFM_ACTIVE": "no", "BLOCKAGE_ACTIVE": "no"})
    return cases.viscous(case, model="constant", mu=0.01)


def test_viscous_alone_equals_batch():
    n = 40
    makers = {"moving": moving, "duct": duct, "inviscid": lambda: cases.bump2d(33, 17, n_iterations=200)}
    alone = {}
    for name, mk in makers.items():
        print("alone", name, flush=True)
        s = Solver.from_case(mk())
        s.advance(n)
        alone[name] = s.state()
        s.close()
    b = Batch()
    members = {name: Solver.from_case(mk()) for name, mk in makers.items()}
    for s in members.values():
        b.admit(s)
    print("batch of", len(b), flush=True)
    b.advance(n // 2)
    b.advance(n - n // 2)
    for name, s in members.items():
        s.status()
        assert np.array_equal(s.state(), alone[name]), "state of %s differs in the batch" % name
    b.close()


def test_turbulence_refused():
    case = cases.bump2d(17, 9)
    case["config"]["TURBULENCE_ACTIVE"] = "yes"
    with pytest.raises(SolverError, match="TURBULENCE_ACTIVE"):
        Solver.from_case(case)


def test_volume_file_of_viscous_case_refused(tmp_path):
    s = Solver.from_case(cases.viscous(cases.bump2d(17, 9)))
    s.advance(2)
    with pytest.raises(SolverError):
        s.write_solution(str(tmp_path / "v.csv"))
    s.close()
