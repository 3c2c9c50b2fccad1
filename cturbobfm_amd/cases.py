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
                 row x in [0.3, 0.5] m with 20 blades at 6000 rpm, camber growing
                 linearly 0 -> 45 deg from the leading edge: metal angle -45 deg at the
                 leading edge (the relative inflow angle at mid-span, atan(-omega r /
                 u_x), is about -46 deg in the body force's sign convention) turning to
                 0 deg at the trailing edge (metal angles 0 -> +45 deg meet the inflow
                 at 90 deg incidence and the solution diverges), blockage ramp
                 1 -> 0.98 across the row; inlet p0 = 101325 Pa, T0 = 288.15 K, axial;
                 radial-equilibrium outlet with INIT_PRESSURE = 95 kPa (the hub pressure
                 the reference freezes, solver.cpp:818-834); periodic k planes (360 deg).
"""
from __future__ import annotations

import math
import os
from typing import Optional, Dict, List, Tuple

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
           time_integration: int = 0, time_step: str = "local", mach_out: float = 0.5,
           outflow: Optional[str] = None) -> Dict:
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
        "OUTFLOW": outflow if outflow else "outlet, %.17g" % p_out,
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
            time_step: str = "local", outlet: str = "radial_equilibrium", off_design: bool = False,
            stall: bool = False, turbo_output: bool = False, outflow: Optional[str] = None) -> Dict:
    """Annular rotor; the last k plane is a bitwise copy of k = 0 (full annulus).
    bfm_model Hall, Lift-Drag or Hall-Thollet (K_n = K_f = K_d = 1, nu = 1.48e-5 as the
    reference's GER4 cases; off_design adds the deviation-pivot term); stall switches on
    STALLED_BFM_ACTIVE with a stall deviation of 0.12 rad.  turbo_output: TURBO_OUTPUT (the
    mass flow, PRtt, TRtt, ETAtt of every iteration; the figures of the current state are
    available either way).  outflow: the OUTFLOW patch string instead of
    "<outlet>, <outlet_pressure>" (e.g. "throttle, k_t" or "mass_flow, target[, p_init]")."""
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
        "deviationAnglePivot": field(np.where(inside, 0.02, 0.0)),
        "deviationAngleStall": field(np.where(inside, 0.12, 0.0)),
    }
    config = {
        "KIND_SOLVER": "Euler", "TOPOLOGY": "3D",
        "FLUID_GAMMA": "1.4", "FLUID_R_CONSTANT": "287.05", "FLUID_MODEL": "Ideal",
        "BFM_ACTIVE": "yes", "BFM_MODEL": bfm_model,
        "BLOCKAGE_ACTIVE": "yes" if blockage else "no",
        "LEADING_EDGE_INDEX": str(le_idx), "TRAILING_EDGE_INDEX": str(te_idx),
        "INFLOW": "inlet, 101325, 288.15, 1, 0, 0",
        "OUTFLOW": outflow if outflow else "%s, %.17g" % (outlet, outlet_pressure),
        "HUB": "inviscid_wall", "SHROUD": "inviscid_wall",
        "PER1": "periodic, 0, 0, 360", "PER2": "periodic, 1, 0, 360",
        "INIT_MACH_NUMBER": "%.17g" % mach, "INIT_TEMPERATURE": "%.17g" % t_init,
        "INIT_PRESSURE": "%.17g" % outlet_pressure, "INIT_DIRECTION": "1, 0, 0",
        "CFL": "%.17g" % cfl, "N_ITERATIONS": str(n_iterations),
        "RESIDUALS_DROP_CONVERGENCE": "16", "SAVE_ITERATIONS_INTERVAL": str(max(n_iterations, 1)),
        "TIME_INTEGRATION_TYPE": str(time_integration), "TIME_STEP_METHOD": time_step,
        "CONVECTION_SCHEME": scheme, "MUSCL_RECONSTRUCTION": "yes" if muscl else "no",
        "FLUX_LIMITER": limiter, "TURBO_OUTPUT": "yes" if turbo_output else "no", "TURBO_DIRECTION": "i",
        "SOLUTION_NAME": "results",
    }
    if time_step == "fixed":
        config["FIXED_TIME_STEP"] = "1e-7"
    if bfm_model == "Hall-Thollet":
        config.update({"HALL_THOLLET_COEFFICIENT_KN": "1.0", "HALL_THOLLET_COEFFICIENT_KF": "1.0",
                       "HALL_THOLLET_COEFFICIENT_KD": "1.0",
                       "HALL_THOLLET_OFF_DESIGN_ACTIVE": "yes" if off_design else "no",
                       "FLUID_KINEMATIC_VISCOSITY": "1.48E-5"})
    if stall:
        config["STALLED_BFM_ACTIVE"] = "yes"
    patches = [("INFLOW", 0, 0, 0, nj, 0, nk), ("OUTFLOW", ni, ni, 0, nj, 0, nk),
               ("HUB", 0, ni, 0, 0, 0, nk), ("SHROUD", 0, ni, nj, nj, 0, nk),
               ("PER1", 0, ni, 0, nj, 0, 0), ("PER2", 0, ni, 0, nj, nk, nk)]
    return {"ni": ni, "nj": nj, "nk": nk, "x": X, "y": Y, "z": Z, "fields": fields,
            "config": config, "patches": patches,
            "name": "rotor3d_%dx%dx%d" % (ni, nj, nk)}


def viscous(case: Dict, model: str = "sutherland", mu: float = 1.8e-5,
            walls: Optional[Dict[str, tuple]] = None) -> Dict:
    """The case with VISCOSITY_ACTIVE = yes: air's Sutherland law (mu_ref 1.716e-5 Pa s at
    273.15 K, S = 110.4 K), or a constant viscosity mu; c_p = 1006 J/(kg K), Pr = 0.708.
    walls: patch name -> wall velocity (m/s) of the walls made no_slip_wall; by default
    every inviscid wall becomes a wall at rest."""
    out = dict(case)
    cfg = dict(case["config"])
    cfg.update({"VISCOSITY_ACTIVE": "yes", "FLUID_CP": "1006", "FLUID_PRANDTL_NUMBER": "0.708"})
    if model == "constant":
        cfg.update({"FLUID_VISCOSITY_MODEL": "constant", "FLUID_MU_CONSTANT": "%.17g" % mu})
    else:
        cfg.update({"FLUID_SUTHERLAND_MU_REF": "1.716E-5", "FLUID_SUTHERLAND_T_REF": "273.15",
                    "FLUID_SUTHERLAND_S": "110.4"})
    if walls is None:
        walls = {k: (0.0, 0.0, 0.0) for k, v in cfg.items() if v == "inviscid_wall"}
    for name, u in walls.items():
        cfg[name] = "no_slip_wall, %.17g, %.17g, %.17g" % tuple(u)
    out["config"] = cfg
    out["name"] = case["name"] + "_viscous"
    return out


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


def read_ini(ini_path: str) -> Dict[str, str]:
    """KEY = value pairs of an input.ini (comment and section lines skipped)."""
    cfg: Dict[str, str] = {}
    with open(ini_path) as f:
        for line in f:
            line = line.strip()
            if not line or line[0] in ";[" or "=" not in line:
                continue
            k, v = line.split("=", 1)
            cfg[k.strip()] = v.strip()
    return cfg


def read_case(ini_path: str) -> Dict:
    """Reads input.ini + grid CSV + boundaries CSV back into a case dict."""
    cfg = read_ini(ini_path)
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


# --------------------------------------------------------------- real-gas table
def covolume_lut(seed: int = 5, n_rho: int = 60, n_e: int = 60, n_pt: int = 40) -> Dict:
    """A seeded smooth analytic CO2-like equation of state tabulated in the reference's
    look-up-table layout (src/fluid_real.cpp:12-209).

    Covolume (Noble-Abel) gas: p = (g-1) rho e / (1 - b rho), e = cv T, with
    R ~ 188.9 J/kg/K, g ~ 1.29, b ~ 9.7e-4 m3/kg (each perturbed by the seed):
      a^2 = g p / (rho (1 - b rho)),  dp/drho|e = (g-1) e / (1 - b rho)^2,
      dp/de|rho = (g-1) rho / (1 - b rho),
      s = cv ln(T/Tr) + R ln((1/rho - b)/(vr - b))   (Tr = 288.15 K, vr at 101325 Pa),
    and the (p, T) and (p, s) tables from the closed-form inverses.  Returns a dict with
    the header values and the three tables (arrays shaped (nx, ny, fields)).
    """
    rng = np.random.default_rng(seed)
    u = rng.uniform(-1.0, 1.0, size=3)
    R = 188.92 * (1.0 + 0.01 * u[0])
    g = 1.29 + 0.02 * u[1]
    b = 9.7e-4 * (1.0 + 0.1 * u[2])
    cv = R / (g - 1.0)
    cp = cv + R
    Tr, pr = 288.15, 101325.0
    vr = R * Tr / pr + b

    def props(rho, e):
        T = e / cv
        p = (g - 1.0) * rho * e / (1.0 - b * rho)
        a = np.sqrt(g * p / (rho * (1.0 - b * rho)))
        s = cv * np.log(T / Tr) + R * np.log((1.0 / rho - b) / (vr - b))
        dpdrho = (g - 1.0) * e / (1.0 - b * rho) ** 2
        dpde = (g - 1.0) * rho / (1.0 - b * rho)
        return p, T, a, s, dpdrho, dpde

    rho_min, rho_max = 0.4, 4.0
    t_min, t_max = 150.0, 450.0
    e_min, e_max = cv * t_min, cv * t_max
    rho, e = np.meshgrid(np.linspace(rho_min, rho_max, n_rho), np.linspace(e_min, e_max, n_e), indexing="ij")
    p, T, a, s, dpdrho, dpde = props(rho, e)
    mu = np.full_like(p, 1.5e-5)
    kappa = np.full_like(p, 0.017)
    rho_e = np.stack([p, T, a, s, mu, kappa, dpdrho, dpde], axis=-1)

    p_min, p_max = 2.0e4, 2.5e5
    pp, tt = np.meshgrid(np.linspace(p_min, p_max, n_pt), np.linspace(t_min, t_max, n_pt), indexing="ij")
    rho_pt = pp / (R * tt + b * pp)
    e_pt = cv * tt
    _, _, a_pt, s_pt, _, _ = props(rho_pt, e_pt)
    p_t = np.stack([rho_pt, e_pt, a_pt, s_pt], axis=-1)

    s_min, s_max = float(s_pt.min()), float(s_pt.max())
    pp, ss = np.meshgrid(np.linspace(p_min, p_max, n_pt), np.linspace(s_min, s_max, n_pt), indexing="ij")
    # s = cp ln T - cv ln Tr + R ln(R / (p (vr - b)))
    T_ps = np.exp((ss + cv * math.log(Tr) - R * np.log(R / (pp * (vr - b)))) / cp)
    rho_ps = pp / (R * T_ps + b * pp)
    e_ps = cv * T_ps
    p_s = np.stack([rho_ps, T_ps, e_ps], axis=-1)
    return {"FLUID_NAME": "CO2_COVOLUME_SEED%d" % seed, "MOLAR_MASS": 8.314462618 / R,
            "GAS_CONSTANT": R, "CRITICAL_PRESSURE": 7.3773e6, "CRITICAL_TEMPERATURE": 304.13,
            "CRITICAL_DENSITY": 467.6, "NOMINAL_GAMMA": g,
            "rho_e": (rho_min, rho_max, e_min, e_max, rho_e),
            "p_t": (p_min, p_max, t_min, t_max, p_t),
            "p_s": (p_min, p_max, s_min, s_max, p_s)}


def write_lut(lut: Dict, path: str) -> str:
    """Writes a table in the reference's .lut text format (fluid_real.cpp:12-209: header
    KEY: value lines, then # SECTION: RHO_E / P_T / P_S blocks whose last bound key
    triggers reading the rows)."""
    def rows(f, arr):
        nx, ny, nf = arr.shape
        for i in range(nx):
            for j in range(ny):
                f.write(" ".join("%.17g" % v for v in arr[i, j]) + "\n")
    with open(path, "w") as f:
        for k in ("FLUID_NAME", "MOLAR_MASS", "GAS_CONSTANT", "CRITICAL_PRESSURE", "CRITICAL_TEMPERATURE",
                  "CRITICAL_DENSITY", "NOMINAL_GAMMA"):
            v = lut[k]
            f.write("%s: %s\n" % (k, v if isinstance(v, str) else "%.17g" % v))
        x0, x1, y0, y1, a = lut["rho_e"]
        f.write("# SECTION: RHO_E\nN_RHO: %d\nN_E: %d\nRHO_MIN: %.17g\nRHO_MAX: %.17g\nE_MIN: %.17g\nE_MAX: %.17g\n"
                % (a.shape[0], a.shape[1], x0, x1, y0, y1))
        rows(f, a)
        x0, x1, y0, y1, a = lut["p_t"]
        f.write("# SECTION: P_T\nN_P: %d\nN_T: %d\nP_MIN: %.17g\nP_MAX: %.17g\nT_MIN: %.17g\nT_MAX: %.17g\n"
                % (a.shape[0], a.shape[1], x0, x1, y0, y1))
        rows(f, a)
        x0, x1, y0, y1, a = lut["p_s"]
        f.write("# SECTION: P_S\nN_P: %d\nN_S: %d\nP_MIN: %.17g\nP_MAX: %.17g\nS_MIN: %.17g\nS_MAX: %.17g\n"
                % (a.shape[0], a.shape[1], x0, x1, y0, y1))
        rows(f, a)
    return path


def with_real_gas(case: Dict, lut_path: str) -> Dict:
    """The same case on the look-up-table fluid (FLUID_MODEL = real, FLUID_TABLE_FILE)."""
    out = dict(case)
    cfg = dict(case["config"])
    cfg["FLUID_MODEL"] = "real"
    cfg["FLUID_TABLE_FILE"] = os.path.abspath(lut_path)
    out["config"] = cfg
    return out
