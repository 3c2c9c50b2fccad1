"""GPU parity of the explicit residual path against the CPU restatement (oracle/).

Bar (FP64 path): per-call residuals and time steps agree to 1e-13 relative (bit-exact
where no libm transcendental enters: the inlet bisection uses pow, the body force asin /
atan2); full runs agree to <= 1e-11 relative in the per-equation log-residual history
and in the final conserved fields, the bar of the FP64 path against the CPU solver.
"""
import numpy as np
import pytest

from cturbobfm_amd import Solver, cases

import oracle as O
import regression_cases as RC

pytestmark = pytest.mark.gpu

RUN_TOL = 1e-11
CALL_TOL = 1e-13


def rel_err(a, b):
    a = np.asarray(a).reshape(5, -1) if np.asarray(a).size % 5 == 0 and np.asarray(a).ndim == 4 else np.asarray(a)
    b = np.asarray(b).reshape(a.shape)
    if a.ndim == 2:
        return max(np.max(np.abs(a[v] - b[v])) / max(np.max(np.abs(b[v])), 1e-300) for v in range(a.shape[0]))
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def perturbed(case, seed=3, amp=0.02):
    oc = O.OracleCase(case)
    return cases.perturb_state(case, oc.state(), amp, seed), oc


CASES = {
    "bump2d_jst": lambda: cases.bump2d(33, 17, n_iterations=200),
    "bump2d_roe_muscl_va": lambda: cases.bump2d(33, 17, n_iterations=200, scheme="ROE", muscl=True,
                                                  limiter="van albada", cfl=1.0),
    "bump2d_roe_muscl_minmod": lambda: cases.bump2d(33, 17, n_iterations=200, scheme="ROE", muscl=True,
                                                      limiter="min mod", cfl=1.0),
    "rotor3d_roe_muscl_minmod_hall": lambda: cases.rotor3d(25, 9, 17, n_iterations=200, scheme="ROE", muscl=True,
                                                            limiter="min mod", cfl=1.2),
    "bump2d_roe_1st_rk3_global": lambda: cases.bump2d(33, 17, n_iterations=200, scheme="ROE", cfl=0.8,
                                                        time_integration=1, time_step="global"),
    "rotor3d_jst_hall": lambda: cases.rotor3d(25, 9, 17, n_iterations=200),
    "rotor3d_roe_muscl_liftdrag": lambda: cases.rotor3d(25, 9, 17, n_iterations=200, scheme="ROE", muscl=True,
                                                          limiter="van leer", bfm_model="Lift-Drag", cfl=1.2),
    "rotor3d_jst_thollet": lambda: cases.rotor3d(25, 9, 17, n_iterations=200, bfm_model="Hall-Thollet"),
    # off-design loss and the stall model: about 160 of the 1071 blade nodes run stalled
    "rotor3d_roe_muscl_thollet_od_stall": lambda: cases.rotor3d(25, 9, 17, n_iterations=200, scheme="ROE", muscl=True,
                                                                  limiter="van albada", bfm_model="Hall-Thollet",
                                                                  off_design=True, stall=True, cfl=1.2),
    # grids smaller than one tile (every stage-kernel tile shape is 16 to 64 nodes wide),
    # and one node beyond a whole number of 28 x 9 tiles in i and j; k pieces of one or two
    # planes
    "rotor3d_jst_tiny_9x4x5": lambda: cases.rotor3d(9, 4, 5, n_iterations=200),
    "rotor3d_roe_muscl_ragged_29x10x5": lambda: cases.rotor3d(29, 10, 5, n_iterations=200, scheme="ROE", muscl=True,
                                                                limiter="van albada", cfl=1.0),
    "bump2d_jst_tiny_5x3": lambda: cases.bump2d(5, 3, n_iterations=200),
    "bump2d_roe_muscl_ragged_30x10": lambda: cases.bump2d(30, 10, n_iterations=200, scheme="ROE", muscl=True,
                                                            limiter="van albada", cfl=1.0),
    # outlets whose hub pressure follows the mass flow every iteration
    "rotor3d_jst_throttle": lambda: cases.rotor3d(25, 9, 17, n_iterations=200, outflow="throttle, 0.01"),
    "rotor3d_jst_massflow": lambda: cases.rotor3d(25, 9, 17, n_iterations=200, outflow="mass_flow, 15"),
    "bump2d_jst_massflow_uniform": lambda: cases.bump2d(33, 17, n_iterations=200, outflow="mass_flow, 60, 90000"),
    "rotor3d_jst_throttle_greitzer": lambda: greitzer(cases.rotor3d(25, 9, 17, n_iterations=200,
                                                                    outflow="throttle, 0.01")),
    # laminar viscous terms: walls at rest or moving (the first stage then reads the state
    # with the wall momentum reset), constant or Sutherland viscosity; the inlet nodes on
    # the walls take the wall's stagnation pressure.  The 3-D cases are the annulus without
    # its blade row: no-slip walls put u = 0 nodes into a blade row, where the body force
    # of the reference (and of both implementations) is not finite.
    "bump2d_jst_viscous": lambda: cases.viscous(cases.bump2d(33, 17, n_iterations=200), model="constant", mu=5e-3),
    "bump2d_roe_rk3_global_viscous_suth_moving": lambda: cases.viscous(
        cases.bump2d(33, 17, n_iterations=200, scheme="ROE", cfl=0.8, time_integration=1, time_step="global"),
        walls={"LOWER_WALL": (40.0, 0.0, 0.0), "UPPER_WALL": (0.0, 0.0, 0.0)}),
    "duct3d_roe_muscl_viscous_moving": lambda: cases.viscous(
        no_bfm(cases.rotor3d(25, 9, 17, n_iterations=200, scheme="ROE", muscl=True, limiter="van albada", cfl=1.2)),
        model="constant", mu=0.02, walls={"HUB": (0.0, 0.0, 0.0), "SHROUD": (0.0, 0.0, 30.0)}),
    "duct3d_jst_viscous_suth": lambda: cases.viscous(no_bfm(cases.rotor3d(25, 9, 17, n_iterations=200))),
    # axisymmetric: the wedge faces enter the gradients (their face vectors do not cancel)
    "ger4_axi_thollet_viscous_suth": lambda: cases.viscous(RC.load_case("GER4_Axisymmetric_HallThollet"), walls={}),
    # bladeless axisymmetric channel with no-slip hub (at rest) and shroud (moving): the
    # wall paths (first-stage wall momentum, wall inlet faces, residual zeroing) on the
    # wedge topology
    "axi_channel_viscous_walls_moving": lambda: cases.viscous(
        RC.load_case("AxisymmetricSwirlingChannel"), model="constant", mu=2e-3,
        walls={"LOWER_WALL": (0.0, 0.0, 0.0), "UPPER_WALL": (0.0, 0.0, 25.0)}),
}


def no_bfm(case):
    case["config"].update({"BFM_ACTIVE": "no", "BLOCKAGE_ACTIVE": "no"})
    return case


def greitzer(case):
    """Greitzer plenum behind the throttle: fixed time step, plenum of 0.05 m^3."""
    case["config"].update({"ENABLE_GREITZER_MODELING": "yes", "GREITZER_PLENUM_VOLUME": "0.05",
                           "TIME_STEP_METHOD": "fixed", "FIXED_TIME_STEP": "2e-6"})
    return case


@pytest.mark.parametrize("name", list(CASES))
def test_geometry_and_timestep(name):
    case = CASES[name]()
    print("case", name, flush=True)
    s = Solver.from_case(case)
    U, oc = perturbed(case)
    for w, ow in ((0, 0), (1, 1), (2, 2), (3, 6)):
        assert np.array_equal(s.geometry(w), oc.geometry(ow)), "geometry array %d differs" % w
    dt_g = s.timestep(U)
    dt_o = oc.timestep(U)
    if "suth" in name:   # Sutherland's law calls pow: libm and the device may differ by an ulp
        assert rel_err(dt_g, dt_o) <= 1e-15
    else:
        assert np.array_equal(dt_g, dt_o), "time step differs: %g" % rel_err(dt_g, dt_o)


@pytest.mark.parametrize("name", list(CASES))
def test_residual_call(name):
    case = CASES[name]()
    print("case", name, flush=True)
    s = Solver.from_case(case)
    U, oc = perturbed(case)
    oc.set_state(U)
    s.set_state(U)
    Ro = oc.residual(U, it=1)
    Rg = s.residual(U, it=1)
    err = rel_err(Rg, Ro)
    exact = np.mean(Rg == Ro)
    print("residual rel err %.3e, bit-exact fraction %.4f" % (err, exact))
    assert err <= CALL_TOL


@pytest.mark.parametrize("name", list(CASES))
def test_run_history_and_fields(name):
    case = CASES[name]()
    print("case", name, flush=True)
    n = 60
    s = Solver.from_case(case)
    oc = O.OracleCase(case)
    lo, conv_o, _ = oc.run(n, threads=8)
    lg = s.advance(n)
    assert lg.shape == lo.shape
    hist_err = np.max(np.abs(lg - lo))
    Ug, Uo = s.state(), oc.state()
    f_err = rel_err(Ug, Uo)
    print("history abs err %.3e, field rel err %.3e" % (hist_err, f_err))
    assert hist_err <= RUN_TOL
    assert f_err <= RUN_TOL
    # body-force side outputs and time step of the last iteration
    fi_g, fv_g, dv_g, dt_g = s.side_outputs()
    fi_o, fv_o, dv_o, dt_o = oc.side_outputs()
    for a, b in ((fi_g, fi_o), (fv_g, fv_o), (dv_g, dv_o), (dt_g, dt_o)):
        if np.max(np.abs(b)) > 0:
            assert np.max(np.abs(a - b)) / np.max(np.abs(b)) <= RUN_TOL


def test_turbo_outputs():
    case = cases.rotor3d(25, 9, 17, n_iterations=100)
    s = Solver.from_case(case)
    oc = O.OracleCase(case)
    oc.run(40, threads=8)
    s.advance(40)
    tg, to = s.turbo(), oc.turbo()
    print("turbo gpu", tg[6:], "oracle", to[6:])
    assert np.max(np.abs(tg - to) / np.maximum(np.abs(to), 1e-30)) <= RUN_TOL


def test_convergence_stop_matches():
    case = cases.bump2d(17, 9, n_iterations=3000)
    case["config"]["RESIDUALS_DROP_CONVERGENCE"] = "3"
    s = Solver.from_case(case)
    oc = O.OracleCase(case)
    lo, conv_o, _ = oc.run(3000, threads=4)
    s.solve(3000)
    print("converged at gpu %d oracle %d" % (s.converged_at, conv_o))
    assert conv_o > 0 and s.converged_at == conv_o


def test_determinism_bitwise():
    case = cases.rotor3d(25, 9, 17, n_iterations=30)
    a, b = Solver.from_case(case), Solver.from_case(case)
    la, lb = a.advance(20), b.advance(20)
    assert np.array_equal(la, lb)
    assert np.array_equal(a.state(), b.state())


def test_thollet_stall_exercised():
    """The stalled-blade branch is taken (camber-normal force direction, no viscous
    force) at nodes whose deviation exceeds the stall deviation, and the GPU agrees with
    the restatement on which nodes those are."""
    case = CASES["rotor3d_roe_muscl_thollet_od_stall"]()
    s = Solver.from_case(case)
    oc = O.OracleCase(case)
    oc.run(60, threads=8)
    s.advance(60)
    dv_g = s.side_outputs()[2]
    dv_o = oc.side_outputs()[2]
    stall = 0.12
    n_stalled = int(np.sum(dv_o > stall))
    print("stalled nodes", n_stalled, "of", int(np.sum(dv_o != 0)), flush=True)
    assert n_stalled > 50
    assert np.array_equal(dv_g > stall, dv_o > stall)


def test_thollet_missing_coefficient():
    case = CASES["rotor3d_jst_thollet"]()
    del case["config"]["HALL_THOLLET_COEFFICIENT_KF"]
    from cturbobfm_amd import SolverError
    with pytest.raises(SolverError, match="HALL_THOLLET_COEFFICIENT_KF"):
        Solver.from_case(case)


def time_dependent(case):
    """ACCELERATION_ACTIVE (rotational speed ramping from 0.5 to 1.0 of its value between
    1 and 4 microseconds) and PERTURBATION_BODY_FORCE (40 % Gaussian about a blade-row
    point, radius 4 cm, between 1 and 4 microseconds) on a fixed 1e-7 s time step."""
    case["config"].update({
        "TIME_STEP_METHOD": "fixed", "FIXED_TIME_STEP": "1e-7",
        "ACCELERATION_ACTIVE": "yes", "ACCELERATION_INITIAL_TIME": "1e-6", "ACCELERATION_FINAL_TIME": "4e-6",
        "ROTATIONAL_SPEED_SCALING_FACTOR_INITIAL": "0.5", "ROTATIONAL_SPEED_SCALING_FACTOR_FINAL": "1.0",
        "PERTURBATION_BODY_FORCE": "yes", "PERTURBATION_CENTER": "0.4, 0.25, 0.0",
        "PERTURBATION_RADIAL_EXTENSION": "0.04", "PERTURBATION_SCALING_FACTOR": "0.4",
        "PERTURBATION_TIME_START": "1e-6", "PERTURBATION_TIME_DURATION": "3e-6"})
    return case


@pytest.mark.parametrize("model", ["Hall", "Hall-Thollet"])
def test_time_dependent_body_force(model):
    """The rotational-speed ramp and the body-force perturbation over 60 iterations (6
    microseconds: before, inside and after both windows), GPU against the oracle at the
    run bar; the physical time both sides see is the sum of the fixed steps."""
    case = time_dependent(cases.rotor3d(25, 9, 17, n_iterations=100, bfm_model=model))
    print("case", case["name"], model, flush=True)
    s = Solver.from_case(case)
    oc = O.OracleCase(case)
    lo, _, _ = oc.run(60, threads=8)
    lg = s.advance(60)
    hist = np.max(np.abs(lg - lo))
    f_err = rel_err(s.state(), oc.state())
    print("history abs err %.3e, field rel err %.3e" % (hist, f_err), flush=True)
    assert hist <= RUN_TOL and f_err <= RUN_TOL
    fi_g, _, dv_g, _ = s.side_outputs()
    fi_o, _, dv_o, _ = oc.side_outputs()
    assert np.max(np.abs(fi_g - fi_o)) / np.max(np.abs(fi_o)) <= RUN_TOL
    # the options matter: the same case without them ends elsewhere
    plain = cases.rotor3d(25, 9, 17, n_iterations=100, bfm_model=model, time_step="fixed")
    p = Solver.from_case(plain)
    p.advance(60)
    assert rel_err(p.state(), s.state()) > 1e-8


def test_acceleration_needs_fixed_time_step():
    from cturbobfm_amd import SolverError
    case = time_dependent(cases.rotor3d(17, 5, 9))
    case["config"]["TIME_STEP_METHOD"] = "local"
    with pytest.raises(SolverError, match="fixed time step"):
        Solver.from_case(case)
