"""The reference's regression cases (CircularBump2D: Roe 2-D; CircularBump3D_PeriodicSwirl:
JST 3-D, 45-degree sector, cylindrical inlet, radial-equilibrium outlet;
AxisymmetricSwirlingChannel: JST axisymmetric wedge; GER4_Axisymmetric_HallThollet and
GER4_FullAnnulus: the GER4 compressor rotor with the Hall-Thollet body force, axisymmetric
and 360-degree periodic with the Gong rotor source) run to the reference's convergence
test with the CPU restatement (Sod_Test_Roe2: 1-D Roe MUSCL shock tube with transparent ends, fixed time step, started
from the reference's restart file, 2500 iterations) and compared with the reference's
reference.csv on
Density, Pressure and Temperature by relative L2, with the reference's own tolerance
3e-2 (test/regression_tests/test_cases.py:82-83)."""
import pytest

import oracle as O
import regression_cases as RC


@pytest.mark.parametrize("name", list(RC.CONFIGS))
def test_oracle_matches_reference_solution(name, tmp_path):
    case = RC.load_case(name)
    oc = O.OracleCase(case)
    U0 = RC.initial_state(name, str(tmp_path / name))
    if U0 is not None:
        oc.set_state(U0)
    lr, conv, _ = oc.run(RC.max_iterations(name), threads=4)
    assert (conv > 0) == RC.converges(name), "convergence differs from the reference's run"
    ref = RC.load_reference_solution(name)
    got = RC.derived_fields(oc.state())
    for key in ("Density", "Pressure", "Temperature"):
        err = RC.rel_l2(ref[key], got[key])
        print(name, key, "%.3e" % err)
        assert err < 3e-2
