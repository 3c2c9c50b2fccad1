"""Synthetic workloads for the explicit residual path, in the reference's input formats.

A case is a plain dict:

    {"ni", "nj", "nk": int,
     "x", "y", "z": float64 arrays shaped (ni, nj, nk),
     "fields": {grid-CSV column name: float64 array (ni, nj, nk)},
     "config": {input.ini key: value string},
     "patches": [(NAME, I_MIN, I_MAX, J_MIN, J_MAX, K_MIN, K_MAX), ...]}

which is exactly what the reference reads from its grid CSV (input.cpp:14-82, columns of
types.hpp:1084-1126), its input.ini (config.cpp:15-44) and its boundaries CSV
(solver.cpp:232-294).  ``write_case`` writes those three files so that the same case can
be run as ``turbobfm input.ini``.

Workloads (sizes are node counts):
  * ``bump2d``   2-D channel 3 x 1 with a 10 %-thick circular-arc bump on x in [1, 2]
                 (arc radius 1.3 centred at (1.5, -1.2)); uniform in x, tanh-clustered to
                 the lower wall; JST, RK41A, CFL 2, local dt; inlet p0 = 101325 Pa,
                 T0 = 288.15 K; outlet static pressure for M = 0.5.
  * ``rotor3d``  full annulus, hub r = 0.20 m, tip r = 0.30 m, x in [0, 0.8] m; blade
                 row x in [0.3, 0.5] m with 20 blades at 6000 rpm, metal angle linear
                 from -45 deg (leading edge) to 0 deg (trailing edge), blockage ramp
                 1 -> 0.98 across the row; inlet p0 = 101325 Pa, T0 = 288.15 K, axial;
                 radial-equilibrium outlet with INIT_PRESSURE = 95 kPa (the hub pressure
                 the reference freezes, solver.cpp:818-834); periodic k planes (360 deg).
"""
from __future__ import annotations

import math
import os
from typing import Dict, List, Tuple

import numpy as np

GAMMA = 1.4
R_GAS = 287.05


def _static_from_total(p0: float, t0: float, mach: float) -> Tuple[float, float]:
    f = 1.0 + 0.5 * (GAMMA - 1.0) * mach * mach
    return p0 * f ** (-GAMMA / (GAMMA - 1.0)), t0 / f


def _tanh_cluster(n: int, beta: float) -> np.ndarray:
    """0..1 spacing clustered towards 0 (the wall)."""
    s = np.linspace(0.0, 1.0, n)
    return 1.0 - np.tanh(beta * (1.0 - s)) / math.tanh(beta)


def bump2d(ni: int = 129, nj: int = 65, n_iterations: int = 2000, scheme: str = "JST",
           cfl: float = 2.0, muscl: bool = False, limiter: str = "None",
           time_integration: int = 0, time_step: str = "local", mach_out: float = 0.5) -> Dict:
    p0, t0 = 101325.0, 288.15
    p_out, t_init = _static_from_total(p0, t0, mach_out)
    xs = np.linspace(0.0, 3.0, ni)
    radius, xc, yc = 1.3, 1.5, 0.1 - 1.3
    yb = np.where((xs >= 1.0) & (xs <= 2.0),
                  np.sqrt(np.maximum(radius * radius - (xs - xc) ** 2, 0.0)) + yc, 0.0)
    yb = np.maximum(yb, 0.0)
    eta = _tanh_cluster(nj, 1.5)
    x = np.repeat(xs[:, None], nj, axis=1)
    y = yb[:, None] + (1.0 - yb[:, None]) * eta[None, :]
    x = x[:, :, None].copy()
    y = y[:, :, None].copy()
    z = np.zeros_like(x)
    config = {
        "KIND_SOLVER": "Euler", "TOPOLOGY": "2D",
        "FLUID_GAMMA": "1.4", "FLUID_R_CONSTANT": "287.05", "FLUID_MODEL": "Ideal",
        "INFLOW": "inlet, 101325, 288.15, 1, 0, 0",
        "OUTFLOW": "outlet, %.17g" % p_out,
        "LOWER_WALL": "inviscid_wall", "UPPER_WALL": "inviscid_wall",
        "INIT_MACH_NUMBER": "%.17g" % mach_out, "INIT_TEMPERATURE": "%.17g" % t_init,
        "INIT_PRESSURE": "%.17g" % p_out, "INIT_DIRECTION": "1, 0, 0",
        "CFL": "%.17g" % cfl, "N_ITERATIONS": str(n_iterations),
        "RESIDUALS_DROP_CONVERGENCE": "16", "SAVE_ITERATIONS_INTERVAL": str(max(n_iterations, 1)),
        "TIME_INTEGRATION_TYPE": str(time_integration), "TIME_STEP_METHOD": time_step,
        "CONVECTION_SCHEME": scheme, "MUSCL_RECONSTRUCTION": "yes" if muscl else "no",
        "FLUX_LIMITER": limiter, "SOLUTION_NAME": "results",
    }
    if time_step == "fixed":
        config["FIXED_TIME_STEP"] = "1e-6"
    patches = [("INFLOW", 0, 0, 0, nj, 0, 1), ("OUTFLOW", ni, ni, 0, nj, 0, 1),
               ("LOWER_WALL", 0, ni, 0, 0, 0, 1), ("UPPER_WALL", 0, ni, nj, nj, 0, 1)]
    return {"ni": ni, "nj": nj, "nk": 1, "x": x, "y": y, "z": z, "fields": {},
            "config": config, "patches": patches, "name": "bump2d_%dx%d" % (ni, nj)}


def rotor3d(ni: int = 193, nj: int = 49, nk: int = 129, n_iterations: int = 2000,
            scheme: str = "JST", bfm_model: str = "Hall", cfl: float = 2.0,
            muscl: bool = False, limiter: str = "None", outlet_pressure: float = 95000.0,
            length: float = 0.8, blockage: bool = True, time_integration: int = 0,
            time_step: str = "local", outlet: str = "radial_equilibrium") -> Dict:
    """Annular rotor; the last k plane is a bitwise copy of k = 0 (full annulus)."""
    p0, t0 = 101325.0, 288.15
    mach = math.sqrt(max(((p0 / outlet_pressure) ** ((GAMMA - 1.0) / GAMMA) - 1.0) * 2.0 / (GAMMA - 1.0), 1e-4))
    _, t_init = _static_from_total(p0, t0, mach)
    xs = np.linspace(0.0, length, ni)
    rs = 0.20 + 0.10 * np.linspace(0.0, 1.0, nj)
    th = np.linspace(0.0, 2.0 * math.pi, nk)
    X = np.broadcast_to(xs[:, None, None], (ni, nj, nk)).copy()
    Rr = np.broadcast_to(rs[None, :, None], (ni, nj, nk))
    C = np.broadcast_to(np.cos(th)[None, None, :], (ni, nj, nk))
    S = np.broadcast_to(np.sin(th)[None, None, :], (ni, nj, nk))
    Y = (Rr * C).copy()
    Z = (Rr * S).copy()
    Y[:, :, -1] = Y[:, :, 0]
    Z[:, :, -1] = Z[:, :, 0]

    x_le, x_te = 0.3, 0.5
    xi = np.clip((xs - x_le) / (x_te - x_le), 0.0, 1.0)
    inside = (xs >= x_le) & (xs <= x_te)
    kappa = np.where(inside, -math.radians(45.0) * (1.0 - xi), 0.0)
    block = np.where(xs < x_le, 1.0, np.where(xs > x_te, 0.98, 1.0 - 0.02 * xi))

    def field(v):
        return np.broadcast_to(np.asarray(v, dtype=np.float64)[:, None, None], (ni, nj, nk)).copy()

    present = inside.astype(np.float64)
    le_idx = int(np.argmax(inside))
    te_idx = int(ni - 1 - np.argmax(inside[::-1]))
    fields = {
        "blockage": field(block),
        "normalAxial": field(np.where(inside, -np.sin(kappa), 0.0)),
        "normalRadial": field(np.zeros(ni)),
        "normalTangential": field(np.where(inside, np.cos(kappa), 0.0)),
        "bladeMetalAngle": field(kappa),
        "bladeLeanAngle": field(np.zeros(ni)),
        "bladeGasPathAngle": field(np.zeros(ni)),
        "rpm": field(np.where(inside, 6000.0, 0.0)),
        "streamwiseLength": field(xs - x_le),
        "bladePresent": field(present),
        "numberBlades": field(np.where(inside, 20.0, 0.0)),
        "liftDragBeta0": field(kappa),
        "liftDragKpEtamax": field(np.where(inside, 0.05, 0.0)),
    }
    config = {
        "KIND_SOLVER": "Euler", "TOPOLOGY": "3D",
        "FLUID_GAMMA": "1.4", "FLUID_R_CONSTANT": "287.05", "FLUID_MODEL": "Ideal",
        "BFM_ACTIVE": "yes", "BFM_MODEL": bfm_model,
        "BLOCKAGE_ACTIVE": "yes" if blockage else "no",
        "LEADING_EDGE_INDEX": str(le_idx), "TRAILING_EDGE_INDEX": str(te_idx),
        "INFLOW": "inlet, 101325, 288.15, 1, 0, 0",
        "OUTFLOW": "%s, %.17g" % (outlet, outlet_pressure),
        "HUB": "inviscid_wall", "SHROUD": "inviscid_wall",
        "PER1": "periodic, 0, 0, 360", "PER2": "periodic, 1, 0, 360",
        "INIT_MACH_NUMBER": "%.17g" % mach, "INIT_TEMPERATURE": "%.17g" % t_init,
        "INIT_PRESSURE": "%.17g" % outlet_pressure, "INIT_DIRECTION": "1, 0, 0",
        "CFL": "%.17g" % cfl, "N_ITERATIONS": str(n_iterations),
        "RESIDUALS_DROP_CONVERGENCE": "16", "SAVE_ITERATIONS_INTERVAL": str(max(n_iterations, 1)),
        "TIME_INTEGRATION_TYPE": str(time_integration), "TIME_STEP_METHOD": time_step,
        "CONVECTION_SCHEME": scheme, "MUSCL_RECONSTRUCTION": "yes" if muscl else "no",
        "FLUX_LIMITER": limiter, "TURBO_OUTPUT": "yes", "TURBO_DIRECTION": "i",
        "SOLUTION_NAME": "results",
    }
    if time_step == "fixed":
        config["FIXED_TIME_STEP"] = "1e-7"
    patches = [("INFLOW", 0, 0, 0, nj, 0, nk), ("OUTFLOW", ni, ni, 0, nj, 0, nk),
               ("HUB", 0, ni, 0, 0, 0, nk), ("SHROUD", 0, ni, nj, nj, 0, nk),
               ("PER1", 0, ni, 0, nj, 0, 0), ("PER2", 0, ni, 0, nj, nk, nk)]
    return {"ni": ni, "nj": nj, "nk": nk, "x": X, "y": Y, "z": Z, "fields": fields,
            "config": config, "patches": patches,
            "name": "rotor3d_%dx%dx%d" % (ni, nj, nk)}


def perturb_state(case: Dict, U: np.ndarray, amplitude: float, seed: int) -> np.ndarray:
    """Adds a smooth seeded perturbation to a conserved state (5, ni, nj, nk)."""
    rng = np.random.default_rng(seed)
    ni, nj, nk = case["ni"], case["nj"], case["nk"]
    ii = np.linspace(0, 1, ni)[:, None, None]
    jj = np.linspace(0, 1, nj)[None, :, None]
    kk = np.linspace(0, 1, nk)[None, None, :]
    out = U.copy()
    for v in range(5):
        a, b, c = rng.uniform(1, 4, 3)
        ph = rng.uniform(0, 2 * math.pi, 3)
        w = np.sin(a * math.pi * ii + ph[0]) * np.sin(b * math.pi * jj + ph[1]) * np.cos(c * math.pi * kk + ph[2])
        out[v] = U[v] * (1.0 + amplitude * w)
    if case["config"].get("PER1", "").startswith("periodic") and nk > 1:
        out[:, :, :, -1] = out[:, :, :, 0]
    return out


def write_case(case: Dict, folder: str) -> str:
    """Writes grid.csv, boundaries.csv and input.ini; returns the input.ini path."""
    os.makedirs(folder, exist_ok=True)
    ni, nj, nk = case["ni"], case["nj"], case["nk"]
    names = ["x", "y", "z"] + list(case["fields"].keys())
    cols = [case["x"], case["y"], case["z"]] + [case["fields"][n] for n in case["fields"]]
    data = np.stack([c.reshape(-1) for c in cols], axis=1)
    grid = os.path.join(folder, "grid.csv")
    with open(grid, "w") as f:
        f.write("NDIMENSIONS=%d\nNI=%d\nNJ=%d\nNK=%d\n" % (3 if nk > 1 else 2, ni, nj, nk))
        f.write(",".join(names) + "\n")
        np.savetxt(f, data, fmt="%.17g", delimiter=",")
    bnd = os.path.join(folder, "boundaries.csv")
    with open(bnd, "w") as f:
        f.write("NI=%d\nNJ=%d\nNK=%d\n" % (ni, nj, nk))
        f.write("PATCH_NAME,I_MIN,I_MAX,J_MIN,J_MAX,K_MIN,K_MAX\n")
        for p in case["patches"]:
            f.write("%s,%d,%d,%d,%d,%d,%d\n" % tuple(p))
    ini = os.path.join(folder, "input.ini")
    cfg = dict(case["config"])
    cfg["GRID_FILE"] = grid
    cfg["BOUNDARY_CONDITIONS_FILEPATH"] = bnd
    with open(ini, "w") as f:
        f.write("[CFD]\n")
        for k, v in cfg.items():
            f.write("%s = %s\n" % (k, v))
    return ini


def read_case(ini_path: str) -> Dict:
    """Reads input.ini + grid CSV + boundaries CSV back into a case dict."""
    cfg: Dict[str, str] = {}
    with open(ini_path) as f:
        for line in f:
            line = line.strip()
            if not line or line[0] in ";[" or "=" not in line:
                continue
            k, v = line.split("=", 1)
            cfg[k.strip()] = v.strip()
    base = os.path.dirname(os.path.abspath(ini_path))

    def resolve(p):
        return p if os.path.isabs(p) else os.path.join(base, p)

    ni = nj = nk = 0
    header: List[str] = []
    skip = 0
    with open(resolve(cfg["GRID_FILE"])) as f:
        for line in f:
            skip += 1
            if line.startswith("NI="):
                ni = int(line[3:])
            elif line.startswith("NJ="):
                nj = int(line[3:])
            elif line.startswith("NK="):
                nk = int(line[3:])
            elif line.startswith("x,y,z"):
                header = line.strip().split(",")
                break
    data = np.loadtxt(resolve(cfg["GRID_FILE"]), delimiter=",", skiprows=skip, ndmin=2)
    arrs = {h: data[:, c].reshape(ni, nj, nk).copy() for c, h in enumerate(header)}
    patches = []
    with open(resolve(cfg["BOUNDARY_CONDITIONS_FILEPATH"])) as f:
        in_data = False
        for line in f:
            if not line.strip() or "=" in line:
                continue
            if "PATCH_NAME" in line:
                in_data = True
                continue
            if in_data:
                t = [s.strip() for s in line.split(",")]
                patches.append((t[0],) + tuple(int(v) for v in t[1:7]))
    fields = {h: a for h, a in arrs.items() if h not in ("x", "y", "z")}
    return {"ni": ni, "nj": nj, "nk": nk, "x": arrs["x"], "y": arrs["y"], "z": arrs["z"],
            "fields": fields, "config": cfg, "patches": patches,
            "name": os.path.basename(base)}
