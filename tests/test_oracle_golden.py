"""Pins the CPU restatement (oracle/) against golden vectors written by the reference's
own translation units (tests/golden/gen_golden.py).  Agreement is bit-exact: both sides
are built without FMA contraction and follow the same operation order."""
import os

import numpy as np
import pytest

from cturbobfm_amd import cases

import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def fixture_case(name):
    if name == "bump2d":
        return cases.bump2d(9, 6)
    if name == "rotor3d_hall":
        return cases.rotor3d(11, 5, 7, bfm_model="Hall")
    if name == "rotor3d_liftdrag":
        return cases.rotor3d(11, 5, 7, bfm_model="Lift-Drag", scheme="ROE", muscl=True, limiter="van albada")
    if name == "rotor3d_thollet":
        return cases.rotor3d(11, 5, 7, bfm_model="Hall-Thollet")
    if name == "rotor3d_thollet_od_stall":
        return cases.rotor3d(11, 5, 7, bfm_model="Hall-Thollet", off_design=True, stall=True)
    raise KeyError(name)


@pytest.mark.parametrize("name", ["bump2d", "rotor3d_hall", "rotor3d_liftdrag"])
def test_geometry(name):
    g = load("geometry_%s.npz" % name)
    oc = O.OracleCase(fixture_case(name))
    for w in (0, 1, 2, 6):
        assert np.array_equal(oc.geometry(w), g["geom%d" % w]), "geometry %d" % w
    if "gradb" in g:
        assert np.array_equal(oc.geometry(8), g["gradb"])


@pytest.mark.parametrize("name", ["bump2d", "rotor3d_hall", "rotor3d_liftdrag"])
def test_face_flux(name):
    f = load("flux_%s.npz" % name)
    oc = O.OracleCase(fixture_case(name))
    for q in range(len(f["F"])):
        out = oc.face_flux(f["Ull"][q], f["Ul"][q], f["Ur"][q], f["Urr"][q], f["S"][q])
        assert np.array_equal(out, f["F"][q]), q


@pytest.mark.parametrize("lim", ["van_albada", "van_leer", "min_mod"])
def test_muscl_limiters(lim):
    f = load("flux_roe_muscl_%s.npz" % lim)
    oc = O.OracleCase(cases.bump2d(9, 6, scheme="ROE", muscl=True, limiter=lim.replace("_", " ")))
    for q in range(len(f["F"])):
        out = oc.face_flux(f["Ull"][q], f["Ul"][q], f["Ur"][q], f["Urr"][q], f["S"][q])
        assert np.array_equal(out, f["F"][q]), q


def test_boundary_fluxes():
    b = load("boundary.npz")
    oc = O.OracleCase(fixture_case("rotor3d_hall"))
    for q in range(len(b["F"])):
        out = oc.boundary_flux(int(b["kind"][q]), b["values"][q], b["U"][q], b["S"][q],
                               b["face"][q].astype(np.int64), int(b["iteration"]), b["profile"])
        assert np.array_equal(out, b["F"][q]), (q, int(b["kind"][q]))


@pytest.mark.parametrize("name", ["rotor3d_hall", "rotor3d_liftdrag", "rotor3d_thollet", "rotor3d_thollet_od_stall"])
def test_bfm_source(name):
    b = load("bfm_%s.npz" % name)
    oc = O.OracleCase(fixture_case(name))
    for q in range(len(b["S"])):
        i, j, k = (int(v) for v in b["idx"][q])
        s, fi, fv, dv = oc.bfm_source(i, j, k, b["W"][q])
        assert np.array_equal(s, b["S"][q]), q
        assert np.array_equal(fi, b["finv"][q]) and np.array_equal(fv, b["fvisc"][q]), q
        assert dv == b["dev"][q], q


def test_fluid_relations():
    f = load("fluid.npz")
    oc = O.OracleCase(fixture_case("rotor3d_hall"))
    for q in range(len(f["W"])):
        assert np.array_equal(oc.fluid_state(f["W"][q]), f["out"][q]), q


def test_radial_equilibrium():
    r = load("radial_eq.npz")
    p = O.radial_equilibrium(r["rho"], r["ut"], r["r"], float(r["hub"]))
    assert np.array_equal(p, r["p"])


def test_against_reference_build_fresh_inputs(ref_build):
    """Where the reference build exists, also compare on fresh random inputs."""
    import ctypes as C
    R = ref_build
    rng = np.random.default_rng(7)
    for case in (cases.bump2d(9, 6, scheme="ROE", muscl=True, limiter="van leer"),
                 cases.rotor3d(11, 5, 7, bfm_model="Hall")):
        ini = cases.write_case(case, os.path.join(os.path.dirname(GOLD), "..", "cases_tmp", "fresh_" + case["name"]))
        assert R.ref_case_load(ini.encode()) == 0
        oc = O.OracleCase(case)
        for _ in range(100):
            Us = []
            for _q in range(4):
                rho = rng.uniform(0.5, 2.0)
                u = rng.uniform(-300, 300, 3)
                p = rng.uniform(3e4, 2e5)
                et = p / (0.4 * rho) + 0.5 * u @ u
                Us.append(np.array([rho, rho * u[0], rho * u[1], rho * u[2], rho * et]))
            S = rng.normal(size=3)
            a = oc.face_flux(*Us, S)
            b = np.zeros(5)
            dp = lambda x: x.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
            R.ref_face_flux(*[dp(u) for u in Us], dp(S), dp(b))
            assert np.array_equal(a, b)
