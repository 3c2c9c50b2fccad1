"""The reference's regression cases (test/regression_tests/<case>/) as case dicts.

Grid and reference-solution files are data copied from the reference's test folders
into tests/golden/regression/; the boundaries CSV the reference expects is not shipped
there (its .gitignore drops *.csv), so the patch ranges are synthesized from the patch
names and grid sizes.  The configuration values are those of each case's input.ini."""
import os

import numpy as np

from cturbobfm_amd import cases

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "regression")

COMMON = {"KIND_SOLVER": "Euler", "FLUID_GAMMA": "1.4", "FLUID_R_CONSTANT": "287.05", "FLUID_MODEL": "Ideal",
          "TIME_INTEGRATION_TYPE": "0", "TIME_STEP_METHOD": "local", "N_ITERATIONS": "99999",
          "SAVE_ITERATIONS_INTERVAL": "250", "LOWER_WALL": "inviscid_wall", "UPPER_WALL": "inviscid_wall",
          "SOLUTION_NAME": "results", "SAVE_UNSTEADY": "no"}

CONFIGS = {
    "CircularBump2D": dict(COMMON, TOPOLOGY="2D", INFLOW="inlet, 140513.23, 316.224, 1, 0, 0",
                           OUTFLOW="outlet, 101300", INIT_MACH_NUMBER="0.5", INIT_TEMPERATURE="285",
                           INIT_PRESSURE="101335", INIT_DIRECTION="1, 0, 0", CFL="0.6",
                           RESIDUALS_DROP_CONVERGENCE="4", CONVECTION_SCHEME="Roe", MUSCL_RECONSTRUCTION="no",
                           FLUX_LIMITER="van leer"),
    "CircularBump3D_PeriodicSwirl": dict(COMMON, TOPOLOGY="3D", INFLOW="inlet, 140513.23, 316.224, 1, 0, -0.5",
                                         OUTFLOW="radial_equilibrium, 101300", INLET_REFERENCE_FRAME="cylindrical",
                                         PER1="periodic, 0, 0, 45", PER2="periodic, 1, 0, 45",
                                         INIT_MACH_NUMBER="0.7", INIT_TEMPERATURE="285", INIT_PRESSURE="101335",
                                         INIT_DIRECTION="1, 0, 0", CFL="1.3", RESIDUALS_DROP_CONVERGENCE="4",
                                         CONVECTION_SCHEME="JST", MUSCL_RECONSTRUCTION="no",
                                         FLUX_LIMITER="van leer"),
    "AxisymmetricSwirlingChannel": dict(COMMON, TOPOLOGY="axisymmetric", INFLOW="inlet, 130000, 300, 1, 0, 0.5",
                                        OUTFLOW="radial_equilibrium, 101300", INIT_MACH_NUMBER="0.5",
                                        INIT_TEMPERATURE="300", INIT_PRESSURE="101300",
                                        INIT_DIRECTION="1, 0, 0.3", CFL="1.2", RESIDUALS_DROP_CONVERGENCE="3",
                                        CONVECTION_SCHEME="JST"),
}

# GER4 compressor rotor with the Hall-Thollet body force (blockage on, on-design); the
# axisymmetric and full-annulus (360-degree periodic) variants share these keys
GER4 = dict(COMMON, BFM_ACTIVE="yes", BLOCKAGE_ACTIVE="yes", BFM_MODEL="Hall-Thollet",
            HALL_THOLLET_COEFFICIENT_KN="1.0", HALL_THOLLET_COEFFICIENT_KF="1.0",
            HALL_THOLLET_COEFFICIENT_KD="1.0", HALL_THOLLET_OFF_DESIGN_ACTIVE="no",
            FLUID_KINEMATIC_VISCOSITY="1.48E-5", FLUID_NAME="Air", INFLOW="inlet, 101300, 288.15, 1, 0, 0",
            OUTFLOW="radial_equilibrium, 100000", INIT_MACH_NUMBER="0.45", INIT_TEMPERATURE="288.15",
            INIT_PRESSURE="101325", INIT_DIRECTION="1, 0, 0", ROTATION_AXIS="1, 0, 0", CFL="1.3",
            RESIDUALS_DROP_CONVERGENCE="4", TURBO_OUTPUT="yes", TURBO_DIRECTION="i", CONVECTION_SCHEME="JST")
CONFIGS["GER4_Axisymmetric_HallThollet"] = dict(GER4, TOPOLOGY="axisymmetric")
CONFIGS["GER4_FullAnnulus"] = dict(GER4, TOPOLOGY="3D", PER1="periodic, 0, 0, 360.0",
                                   PER2="periodic, 1, 0, 360.0")
# Sod shock tube, 1-D Roe + MUSCL van Leer, transparent ends, fixed time step, started
# from the reference's restart file (2500 iterations; the 16-order drop is never reached)
CONFIGS["Sod_Test_Roe2"] = dict(COMMON, TOPOLOGY="1D", FLUID_NAME="Air", LEFT_END="transparent",
                                RIGHT_END="transparent", CFL="0.6", N_ITERATIONS="2500",
                                RESIDUALS_DROP_CONVERGENCE="16", TIME_STEP_METHOD="fixed",
                                FIXED_TIME_STEP="0.0001", CONVECTION_SCHEME="Roe", MUSCL_RECONSTRUCTION="yes",
                                FLUX_LIMITER="van leer", RESTART_SOLUTION="yes",
                                RESTART_SOLUTION_FILEPATH=os.path.join(HERE, "Sod_Test_Roe2",
                                                                       "grid_250_RestartData.csv"))
GRID_FILES = {"Sod_Test_Roe2": "grid_250_OnlyCoords.csv"}


def max_iterations(name):
    """Iteration limit of a case: N_ITERATIONS when the case runs a fixed count."""
    return min(int(float(CONFIGS[name]["N_ITERATIONS"])), 40000)


def converges(name):
    return float(CONFIGS[name]["RESIDUALS_DROP_CONVERGENCE"]) < 16


def load_case(name):
    folder = os.path.join(HERE, name)
    cfg = dict(CONFIGS[name])
    grid = os.path.join(folder, GRID_FILES.get(name, "grid.csv"))
    ni = nj = nk = 0
    skip = 0
    with open(grid) as f:
        for line in f:
            skip += 1
            if line.startswith("NI="):
                ni = int(line[3:])
            elif line.startswith("NJ="):
                nj = int(line[3:])
            elif line.startswith("NK="):
                nk = int(line[3:])
            elif line.startswith("x,y,z"):
                break
        header = line.strip().split(",")
    data = np.loadtxt(grid, delimiter=",", skiprows=skip, ndmin=2)
    X, Y, Z = (data[:, c].reshape(ni, nj, nk).copy() for c in range(3))
    # body-force input fields (blockage, camber normal, rpm, ...) after x, y, z
    fields = {h: data[:, c].reshape(ni, nj, nk).copy() for c, h in enumerate(header) if c >= 3}
    patches = [("INFLOW", 0, 0, 0, nj, 0, nk), ("OUTFLOW", ni, ni, 0, nj, 0, nk),
               ("LOWER_WALL", 0, ni, 0, 0, 0, nk), ("UPPER_WALL", 0, ni, nj, nj, 0, nk)]
    if cfg["TOPOLOGY"] == "1D":
        patches = [("LEFT_END", 0, 0, 0, nj, 0, nk), ("RIGHT_END", ni, ni, 0, nj, 0, nk)]
    elif nk > 1:
        patches += [("PER1", 0, ni, 0, nj, 0, 0), ("PER2", 0, ni, 0, nj, nk, nk)]
    return {"ni": ni, "nj": nj, "nk": nk, "x": X, "y": Y, "z": Z, "fields": fields, "config": cfg,
            "patches": patches, "name": name}


def load_reference_solution(name):
    path = os.path.join(HERE, name, "reference.csv")
    with open(path) as f:
        dims = [int(f.readline().split("=")[1]) for _ in range(3)]
        header = f.readline().strip().split(",")
    data = np.loadtxt(path, delimiter=",", skiprows=4, ndmin=2)
    return {h: data[:, c].reshape(dims) for c, h in enumerate(header)}


def derived_fields(U, gamma=1.4, R=287.05):
    rho = U[0]
    u2 = (U[1] ** 2 + U[2] ** 2 + U[3] ** 2) / rho ** 2
    p = (gamma - 1) * rho * (U[4] / rho - 0.5 * u2)
    return {"Density": rho, "Pressure": p, "Temperature": p / (rho * R)}


def rel_l2(ref, res):
    return float(np.linalg.norm(ref - res) / np.linalg.norm(ref))


def case_files(name, folder):
    """input.ini + grid + boundaries for `turbobfm input.ini` runs."""
    return cases.write_case(load_case(name), folder)


def initial_state(name, folder):
    """The case's initial state: the restart file through the product's host reader
    (pinned on its own in test_io.py) when RESTART_SOLUTION is set, else None (the
    uniform INIT_* state every solver builds itself)."""
    if CONFIGS[name].get("RESTART_SOLUTION") != "yes":
        return None
    import cturbobfm_amd as T
    return T.initial_state_from_ini(case_files(name, folder))[0]
