"""Look-up-table fluid (FLUID_MODEL = real) in the CPU restatement.

Pinned two ways:
  * golden vectors written by the reference's own FluidReal and the path classes on it
    (tests/golden/lut.npz, from tests/golden/gen_golden.py; the seeded covolume-gas table
    is stored in the fixture): the 17 point functions, JST and Roe-MUSCL face fluxes,
    inlet / outlet / radial-equilibrium / wall fluxes, p a h_t T p_t T_t s, and the
    initial state -- bit-exact;
  * what the mathematics fixes: Table2D reproduces a bilinear function, clamps outside
    its range, returns node values at the nodes, and the e(p, rho) bisection inverts
    p(rho, e) to its stated tolerance; the tabulated covolume gas matches its closed form.
"""
import os

import numpy as np
import pytest

from cturbobfm_amd import cases

import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def gold():
    return np.load(os.path.join(GOLD, "lut.npz"), allow_pickle=False)


def lut_from_fixture(g):
    b = g["bounds"]
    R, gam, mm = (float(v) for v in g["header"])
    return {"FLUID_NAME": "CO2_COVOLUME_SEED5", "MOLAR_MASS": mm, "GAS_CONSTANT": R,
            "CRITICAL_PRESSURE": 7.3773e6, "CRITICAL_TEMPERATURE": 304.13, "CRITICAL_DENSITY": 467.6,
            "NOMINAL_GAMMA": gam,
            "rho_e": (*b[0], g["rho_e"]), "p_t": (*b[1], g["p_t"]), "p_s": (*b[2], g["p_s"])}


@pytest.fixture(scope="module")
def lut_path(gold, tmp_path_factory):
    return cases.write_lut(lut_from_fixture(gold), str(tmp_path_factory.mktemp("lut") / "covolume.lut"))


def real_case(lut_path, **kw):
    return cases.with_real_gas(cases.rotor3d(11, 5, 7, **kw), lut_path)


def test_point_functions(gold, lut_path):
    oc = O.OracleCase(real_case(lut_path))
    pts = gold["points"]
    for which in range(pts.shape[0]):
        for x, y, ref in pts[which]:
            assert oc.lut_eval(which, x, y) == ref, (which, x, y)


@pytest.mark.parametrize("label,kw", [("jst", {}),
                                      ("roe", {"scheme": "ROE", "muscl": True, "limiter": "van albada"})])
def test_face_flux(gold, lut_path, label, kw):
    oc = O.OracleCase(real_case(lut_path, **kw))
    U, S, F = gold["flux_%s_U" % label], gold["flux_%s_S" % label], gold["flux_%s_F" % label]
    for q in range(F.shape[0]):
        out = oc.face_flux(U[0][q], U[1][q], U[2][q], U[3][q], S[q])
        assert np.array_equal(out, F[q]), q


def test_boundary_fluxes(gold, lut_path):
    oc = O.OracleCase(real_case(lut_path))
    for q in range(len(gold["bnd_F"])):
        out = oc.boundary_flux(int(gold["bnd_kind"][q]), gold["bnd_values"][q], gold["bnd_U"][q],
                               gold["bnd_S"][q], gold["bnd_face"][q].astype(np.int64),
                               int(gold["bnd_iteration"]), gold["bnd_profile"])
        assert np.array_equal(out, gold["bnd_F"][q]), (q, int(gold["bnd_kind"][q]))


def test_fluid_relations(gold, lut_path):
    oc = O.OracleCase(real_case(lut_path))
    for q in range(len(gold["fluid_W"])):
        assert np.array_equal(oc.fluid_state(gold["fluid_W"][q]), gold["fluid_out"][q]), q


def test_initial_state(gold, lut_path):
    case = real_case(lut_path)
    case["config"].update({"INIT_MACH_NUMBER": "0.4", "INIT_TEMPERATURE": "290", "INIT_PRESSURE": "95000",
                           "INIT_DIRECTION": "1, 0.2, 0"})
    U = O.OracleCase(case).state()
    assert np.array_equal(U[:, 3, 2, 1], gold["init"])


# ----------------------------------------------------------- closed-form checks
def test_table_bilinear_clamp_and_nodes(tmp_path):
    """A table filled with f = a + b x + c y + d x y is reproduced by the bilinear
    interpolation (to rounding), clamps outside its bounds, and hits its nodes."""
    lut = cases.covolume_lut(seed=2)
    x0, x1, y0, y1, arr = lut["rho_e"]
    nx, ny = arr.shape[:2]
    xs, ys = np.linspace(x0, x1, nx), np.linspace(y0, y1, ny)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    f = 3.0 + 2.0 * X - 1.5e-5 * Y + 4.0e-6 * X * Y
    arr = arr.copy()
    arr[..., 0] = f
    lut["rho_e"] = (x0, x1, y0, y1, arr)
    path = cases.write_lut(lut, str(tmp_path / "bilinear.lut"))
    oc = O.OracleCase(cases.with_real_gas(cases.bump2d(9, 6), path))
    rng = np.random.default_rng(4)
    for _ in range(300):
        x, y = rng.uniform(x0, x1), rng.uniform(y0, y1)
        exact = 3.0 + 2.0 * x - 1.5e-5 * y + 4.0e-6 * x * y
        assert abs(oc.lut_eval(0, x, y) - exact) <= 1e-12 * abs(exact)
        # clamping: far outside equals the nearest boundary point
        assert oc.lut_eval(0, x1 + 10.0, y) == oc.lut_eval(0, x1, y)
        assert oc.lut_eval(0, x, y0 - 1e6) == oc.lut_eval(0, x, y0)
    for i in (0, 7, nx - 1):
        for j in (0, 11, ny - 1):
            assert abs(oc.lut_eval(0, xs[i], ys[j]) - f[i, j]) <= 1e-13 * abs(f[i, j])


def test_bisections_and_closed_form(lut_path):
    oc = O.OracleCase(real_case(lut_path))
    rng = np.random.default_rng(6)
    for _ in range(200):
        rho, e = rng.uniform(0.6, 3.5), rng.uniform(1.0e5, 2.7e5)
        p = oc.lut_eval(0, rho, e)
        e2 = oc.lut_eval(6, p, rho)                     # e(p, rho) by bisection
        assert abs(oc.lut_eval(0, rho, e2) - p) <= 1e-4 * p
    # the tabulated covolume gas against its closed form (interpolation error only)
    lut = cases.covolume_lut(seed=5)
    R, g = lut["GAS_CONSTANT"], lut["NOMINAL_GAMMA"]
    b = 9.7e-4 * (1.0 + 0.1 * np.random.default_rng(5).uniform(-1.0, 1.0, size=3)[2])
    for _ in range(200):
        rho, e = rng.uniform(0.6, 3.5), rng.uniform(1.0e5, 2.7e5)
        p_exact = (g - 1.0) * rho * e / (1.0 - b * rho)
        assert abs(oc.lut_eval(0, rho, e) - p_exact) <= 1e-5 * p_exact
