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


def test_full_annulus_residue_sits_on_the_seam_and_hub():
    """Where the full-annulus case differs from reference.csv.  Every input of the case
    (grid radius, blade fields, boundary values) is the same on every k plane, yet the
    reference's own solution is not axisymmetric: its temperature varies by more than 10 %
    over k on the hub row downstream of the rotor (the seam planes 0 and NK-1 carry the
    extremes) and by more than 5 % elsewhere.
    The restatement converges to a seam artifact of the same kind (insensitive to a 1-ulp
    change of the start state or a 1 % change of CFL), and away from the hub row and the
    seam it agrees with reference.csv at the level of the other regression cases."""
    import numpy as np
    name = "GER4_FullAnnulus"
    case = RC.load_case(name)
    for f, a in case["fields"].items():
        assert np.array_equal(a, np.broadcast_to(a[:, :, :1], a.shape)), f
    r = np.sqrt(case["y"] ** 2 + case["z"] ** 2)
    assert np.max(r.max(axis=2) - r.min(axis=2)) < 1e-12
    ref = RC.load_reference_solution(name)
    T = ref["Temperature"]
    hub_spread = float(np.max((T[:, 0].max(axis=1) - T[:, 0].min(axis=1)) / T[:, 0].mean(axis=1)))
    mid_spread = float(np.max((T[:, 1:].max(axis=2) - T[:, 1:].min(axis=2)) / T[:, 1:].mean(axis=2)))
    print("reference.csv temperature spread over k: hub row %.3f, other rows %.3f" % (hub_spread, mid_spread))
    assert hub_spread > 0.1 and mid_spread > 0.05
    oc = O.OracleCase(case)
    _, conv, _ = oc.run(RC.max_iterations(name), threads=4)
    assert conv > 0
    got = RC.derived_fields(oc.state())
    Tg = got["Temperature"]
    assert float(np.max((Tg[:, 0].max(axis=1) - Tg[:, 0].min(axis=1)) / Tg[:, 0].mean(axis=1))) > 0.1
    for key in ("Density", "Pressure", "Temperature"):
        whole = RC.rel_l2(ref[key], got[key])
        core = RC.rel_l2(ref[key][:, 1:, 1:-1], got[key][:, 1:, 1:-1])
        print("%s: relative L2 %.4f, off the hub row and the seam %.4f" % (key, whole, core))
        assert core < 0.012 and core < whole
