"""Closed forms and invariants that pin the parts of the CPU restatement the reference
build cannot exercise (src/solver.cpp needs Eigen and is not built here): residual
assembly, time step, RK stage loop, periodic closure, norms and convergence test.
Each check is chosen so that a plausible mistake (a dropped term, a wrong sign or
index, a transposed operand) fails it."""
import math

import numpy as np
import pytest

from cturbobfm_amd import cases

import oracle as O


def cartesian(ni, nj, nk, dx=1.0, dy=1.0, dz=1.0, topo="3D", patches=None, extra=None):
    xs, ys, zs = np.arange(ni) * dx, np.arange(nj) * dy, np.arange(nk) * dz
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    cfg = {"TOPOLOGY": topo, "FLUID_GAMMA": "1.4", "FLUID_R_CONSTANT": "287.05", "CONVECTION_SCHEME": "JST",
           "TIME_STEP_METHOD": "local", "CFL": "1.5", "INIT_MACH_NUMBER": "0.4", "INIT_TEMPERATURE": "300",
           "INIT_PRESSURE": "1e5", "INIT_DIRECTION": "1, 0.5, 0.25", "WALL": "inviscid_wall",
           "N_ITERATIONS": "10"}
    if extra:
        cfg.update(extra)
    if patches is None:
        patches = [("WALL", 0, 0, 0, nj, 0, nk), ("WALL", ni, ni, 0, nj, 0, nk), ("WALL", 0, ni, 0, 0, 0, nk),
                   ("WALL", 0, ni, nj, nj, 0, nk)]
        if topo == "3D":
            patches += [("WALL", 0, ni, 0, nj, 0, 0), ("WALL", 0, ni, 0, nj, nk, nk)]
    return {"ni": ni, "nj": nj, "nk": nk, "x": X, "y": Y, "z": Z, "fields": {}, "config": cfg,
            "patches": patches, "name": "cart"}


def test_dual_volumes_closed_form():
    """Unit Cartesian grids: dual cells of volume 1 inside, 1/2 on faces, 1/4 on edges,
    1/8 at corners (3-D); 2-D with unit thickness: 1, 1/2, 1/4 (test_mesh_2d/3d KATs)."""
    oc = O.OracleCase(cartesian(6, 6, 6))
    V = oc.geometry(6)
    nb = (np.add.outer(np.add.outer(np.r_[1, 0, 0, 0, 0, 1], np.r_[1, 0, 0, 0, 0, 1]), np.r_[1, 0, 0, 0, 0, 1]))
    assert np.array_equal(V, 0.5 ** nb)
    oc2 = O.OracleCase(cartesian(6, 6, 1, topo="2D"))
    V2 = oc2.geometry(6)[:, :, 0]
    nb2 = np.add.outer(np.r_[1, 0, 0, 0, 0, 1], np.r_[1, 0, 0, 0, 0, 1])
    assert np.array_equal(V2, 0.5 ** nb2)
    # face vectors of interior I faces are (1, 0, 0) * area 1
    SI = oc.geometry(0)
    assert np.array_equal(SI[2, 2, 2], [1.0, 0.0, 0.0])


def test_timestep_closed_form():
    """Interior nodes of a Cartesian grid: dt = CFL / sum_d (|u_d| + a) / h_d."""
    dx, dy, dz = 0.1, 0.2, 0.05
    case = cartesian(7, 6, 5, dx, dy, dz)
    oc = O.OracleCase(case)
    U = oc.state()
    dt = oc.timestep(U)
    rho = U[0, 0, 0, 0]
    u = U[1:4, 0, 0, 0] / rho
    et = U[4, 0, 0, 0] / rho
    p = 0.4 * rho * (et - 0.5 * u @ u)
    a = math.sqrt(1.4 * p / rho)
    lam = (abs(u[0]) + a) / dx + (abs(u[1]) + a) / dy + (abs(u[2]) + a) / dz
    assert abs(dt[3, 3, 2] - 1.5 / lam) <= 1e-14 * dt[3, 3, 2]
    # global method: the minimum everywhere
    case["config"]["TIME_STEP_METHOD"] = "global"
    og = O.OracleCase(case)
    dtg = og.timestep(U)
    assert np.all(dtg == dt.min())


def test_free_stream_preservation():
    """A uniform state on a curved grid gives zero residual at interior nodes (closed
    dual cells, central flux of a constant, zero JST/Roe dissipation)."""
    for scheme in ("JST", "ROE"):
        case = cases.bump2d(21, 11, scheme=scheme)
        oc = O.OracleCase(case)
        U = oc.state()
        R = oc.residual(U)
        flux_scale = U[1].max() ** 2 / U[0].max()
        assert np.max(np.abs(R[:, 2:-2, 2:-2])) <= 1e-11 * flux_scale, scheme


def test_closed_box_conserves_mass_and_energy():
    """Walls on every side: interior faces telescope, walls carry no mass or energy, so
    the residual summed over all nodes vanishes for rho and rho E."""
    case = cartesian(9, 7, 6)
    oc = O.OracleCase(case)
    U = cases.perturb_state(case, oc.state(), 0.05, 11)
    R = oc.residual(U)
    for v in (0, 4):
        assert abs(R[v].sum()) <= 1e-12 * np.abs(R[v]).sum()
    assert np.abs(R[0]).sum() > 0  # the perturbation does produce fluxes


def test_rk_stage_loop_and_norm():
    """One iteration equals the explicit RK41A recursion U_s = U0 - a_s dt/V R(U_{s-1})
    with a = (1/4, 5/14, 14/25, 1) (config.cpp:400-411), and the reported norm is
    log10(||R_last||_2 / N) (solver.cpp:1488-1508)."""
    case = cartesian(8, 7, 6)
    oc = O.OracleCase(case)
    U0 = cases.perturb_state(case, oc.state(), 0.05, 5)
    oc.set_state(U0)
    dt = oc.timestep(U0)
    V = oc.geometry(6)
    U = U0.copy()
    for a in (1.0 / 4.0, 5.0 / 14.0, 14.0 / 25.0, 1.0):
        R = oc.residual(U)
        U = U0 - R * (a * dt / V)
    lr, _, _ = oc.run(1)
    assert np.array_equal(oc.state(), U)
    N = U0[0].size
    expect = np.log10(np.sqrt((R.reshape(5, -1) ** 2).sum(axis=1)) / N)
    assert np.max(np.abs(lr[0] - expect)) <= 1e-13


def test_rk31b_coefficients():
    case = cartesian(7, 6, 5, extra={"TIME_INTEGRATION_TYPE": "1"})
    oc = O.OracleCase(case)
    U0 = cases.perturb_state(case, oc.state(), 0.05, 8)
    oc.set_state(U0)
    dt, V = oc.timestep(U0), oc.geometry(6)
    U = U0.copy()
    for a in (8.0 / 17.0, 17.0 / 20.0, 1.0):
        U = U0 - oc.residual(U) * (a * dt / V)
    oc.run(1)
    assert np.array_equal(oc.state(), U)


def test_full_annulus_node_mates_stay_identical():
    """360 degrees: the node-mate planes k = 0 and k = NK-1 are the same points and the
    periodic closure keeps their states bitwise equal (solver.cpp:1836-1853, 2308-2327)."""
    case = cases.rotor3d(17, 6, 9)
    oc = O.OracleCase(case)
    oc.set_state(cases.perturb_state(case, oc.state(), 0.02, 2))
    oc.run(5)
    U = oc.state()
    assert np.array_equal(U[:, :, :, 0], U[:, :, :, -1])


def test_sector_periodicity_rotates_momentum():
    """A 45-degree sector: after the closure the k = NK-1 plane is the k = 0 plane
    rotated by +45 degrees about x."""
    ni, nj, nk = 9, 5, 7
    ang = math.radians(45.0)
    xs, rs, th = np.linspace(0, 0.4, ni), np.linspace(0.2, 0.3, nj), np.linspace(0, ang, nk)
    X, Rr, T = np.meshgrid(xs, rs, th, indexing="ij")
    case = cases.rotor3d(ni, nj, nk)
    case.update({"x": X, "y": Rr * np.cos(T), "z": Rr * np.sin(T), "fields": {}})
    case["config"].update({"BFM_ACTIVE": "no", "PER1": "periodic, 0, 0, 45", "PER2": "periodic, 1, 0, 45",
                           "OUTFLOW": "outlet, 95000"})
    oc = O.OracleCase(case)
    oc.run(3)
    U = oc.state()
    c, s = math.cos(ang), math.sin(ang)
    y0, z0 = U[2, :, :, 0], U[3, :, :, 0]
    assert np.array_equal(U[0, :, :, -1], U[0, :, :, 0])
    assert np.array_equal(U[2, :, :, -1], y0 * c - z0 * s)
    assert np.array_equal(U[3, :, :, -1], y0 * s + z0 * c)


def test_convergence_test_semantics():
    """Converged iteration: every relevant equation is below the max of iterations 1-5
    minus the drop; rho*w is skipped in 2-D (solver.cpp:2985-3015)."""
    case = cases.bump2d(17, 9, n_iterations=4000)
    case["config"]["RESIDUALS_DROP_CONVERGENCE"] = "2"
    oc = O.OracleCase(case)
    lr, conv, _ = oc.run(4000)
    assert conv > 5
    base = lr[:5].max(axis=0)
    rel = [0, 1, 2, 4]
    assert np.all(lr[conv - 1, rel] < base[rel] - 2)
    assert not np.all(lr[conv - 2, rel] < base[rel] - 2)


def test_radial_equilibrium_free_vortex():
    """Free vortex u_theta = C / r, constant rho: p(r) = p_hub + rho C^2 / 2 (1/r_h^2 - 1/r^2);
    the trapezoid rule converges at second order."""
    errs = []
    for n in (51, 101):
        r = np.linspace(0.2, 0.3, n)
        rho = np.full(n, 1.2)
        Cc = 30.0
        p = O.radial_equilibrium(rho, Cc / r, r, 95000.0)
        exact = 95000.0 + 1.2 * Cc * Cc / 2.0 * (1 / 0.2 ** 2 - 1 / r ** 2)
        errs.append(np.max(np.abs(p - exact)))
    assert errs[1] < errs[0] / 3.5


def test_roe_consistency_and_rotation():
    """Roe: F(U,U) equals the physical flux; rotating velocity and face vector about x
    rotates the momentum flux (frame invariance of the triad construction)."""
    oc = O.OracleCase(cases.bump2d(9, 6, scheme="ROE"))
    rng = np.random.default_rng(3)
    for _ in range(50):
        rho, u, p = rng.uniform(0.8, 1.3), rng.uniform(-100, 100, 3), rng.uniform(8e4, 1.2e5)
        et = p / (0.4 * rho) + 0.5 * u @ u
        U = np.array([rho, *(rho * u), rho * et])
        S = rng.normal(size=3)
        F = oc.face_flux(U, U, U, U, S)
        n = S / np.linalg.norm(S)
        un = u @ n
        ht = et + p / rho
        phys = np.array([rho * un, *(rho * un * u + p * n), rho * un * ht])
        assert np.allclose(F, phys, rtol=1e-12, atol=1e-9 * np.abs(phys).max())
        th = rng.uniform(0, 2 * math.pi)
        Rm = np.array([[1, 0, 0], [0, math.cos(th), -math.sin(th)], [0, math.sin(th), math.cos(th)]])
        U2 = U.copy()
        U2[1:4] = Rm @ U[1:4]
        U3 = U.copy()
        U3[1:4] = Rm @ U[1:4] * 1.01
        U3[0] *= 1.01
        U3[4] *= 1.01
        Ua = U.copy()
        Ua[1:4] *= 1.01
        Ua[0] *= 1.01
        Ua[4] *= 1.01
        Fa = oc.face_flux(U, U, Ua, Ua, S)
        Fb = oc.face_flux(U2, U2, U3, U3, Rm @ S)
        assert np.allclose(Fb[[0, 4]], Fa[[0, 4]], rtol=1e-9)
        assert np.allclose(Fb[1:4], Rm @ Fa[1:4], rtol=1e-9, atol=1e-7 * np.abs(Fa).max())


def test_hall_force_vanishes_on_camber():
    """Hall model: relative velocity lying in the camber surface gives delta = 0 and no
    force (source_bfm_hall.cpp:3-34)."""
    case = cases.rotor3d(11, 5, 7, bfm_model="Hall", blockage=False)
    oc = O.OracleCase(case)
    i = int(np.argmax(case["fields"]["bladePresent"][:, 0, 0] > 0)) + 1
    j, k = 2, 0   # theta = 0: cylindrical (x, r, theta) = cartesian (x, y, z)
    nax = case["fields"]["normalAxial"][i, j, k]
    ntan = case["fields"]["normalTangential"][i, j, k]
    r = math.hypot(case["y"][i, j, k], case["z"][i, j, k])
    om = 6000 * 2 * math.pi / 60
    wx = 120.0
    wt = -nax * wx / ntan          # n . W = 0
    ut = wt + om * r
    W = np.array([1.1, wx, 0.0, ut, 2.5e5])
    s, fi, fv, dev = oc.bfm_source(i, j, k, W)
    assert abs(dev) < 1e-12
    assert np.max(np.abs(s)) < 1e-6 * 1.1 * wx * wx
