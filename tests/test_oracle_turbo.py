"""Turbomachinery figures of the CPU restatement (Solver::updateMassFlows and
updateTurboPerformance, src/solver.cpp:1632-1781): boundary mass flows as sums of
rho u . S over the boundary faces, mass-weighted total pressure and total temperature on
the inlet and outlet boundaries of the turbo direction, their ratios and the total-to-total
efficiency.  Pinned by closed forms on an annulus whose inlet and outlet are the planes
x = 0 and x = L (boundary face vectors normal to x; the face vectors themselves are
pinned against the reference's Mesh in test_oracle_golden.py): a uniform flow carries
rho u_x times the boundary plane's area (the sum of its face vectors), and planes of
uniform state give ratios of the plane values (total pressure and temperature of the
ideal gas in closed form, efficiency (PR^((g-1)/g) - 1) / (TR - 1) for a compression)."""
import numpy as np
import pytest

from cturbobfm_amd import cases

import oracle as O

G, R = 1.4, 287.05


def state(rho, u, p):
    e = p / ((G - 1.0) * rho) + 0.5 * (u[0] ** 2 + u[1] ** 2 + u[2] ** 2)
    return np.array([rho, rho * u[0], rho * u[1], rho * u[2], rho * e])


def totals(rho, u, p):
    m2 = (u[0] ** 2 + u[1] ** 2 + u[2] ** 2) / (G * p / rho)
    f = 1.0 + 0.5 * (G - 1.0) * m2
    return p * f ** (G / (G - 1.0)), p / (rho * R) * f


def annulus(ni=13, nj=6, nk=17):
    """Case, oracle, and the areas of the inlet and outlet planes."""
    case = cases.rotor3d(ni, nj, nk)
    oc = O.OracleCase(case)
    S = oc.geometry(0)                       # I-face vectors [ni+1][nj][nk][3]
    for plane in (S[0], S[-1]):
        assert np.all(plane[..., 1:] == 0.0)  # planar boundaries normal to x
    return case, oc, abs(S[0, ..., 0].sum()), abs(S[-1, ..., 0].sum())


def test_uniform_flow_mass_flow_is_density_velocity_area():
    case, oc, a_in, a_out = annulus()
    U = oc.state()
    rho, u, p = 1.13, (131.0, 17.0, -42.0), 98000.0
    U[...] = state(rho, u, p)[:, None, None, None]
    t = oc.turbo(U)
    for b, a in ((0, a_in), (1, a_out)):    # I_START and I_END boundaries
        assert abs(abs(t[b]) - rho * u[0] * a) <= 1e-13 * rho * u[0] * a
    expect = 0.5 * (rho * u[0] * a_in + rho * u[0] * a_out)
    assert abs(t[6] - expect) <= 1e-13 * expect          # 360 degrees: no scaling
    assert abs(t[7] - 1.0) < 1e-14 and abs(t[8] - 1.0) < 1e-14
    assert t[9] == 1.0                                    # no temperature rise


@pytest.mark.parametrize("pfac, rfac", [(1.3, 1.1), (1.15, 1.0), (0.8, 0.9)])
def test_planes_of_uniform_state_give_the_ratios_of_their_totals(pfac, rfac):
    case, oc, a_in, a_out = annulus()
    U = oc.state()
    rho1, u1, p1 = 1.2, (120.0, 5.0, 30.0), 101000.0
    rho2, p2 = rho1 * rfac, p1 * pfac
    u2 = (u1[0] / rfac, 40.0, -10.0)                      # same mass flux rho u_x
    U[...] = state(rho1, u1, p1)[:, None, None, None]
    U[:, -1] = state(rho2, u2, p2)[:, None, None]
    t = oc.turbo(U)
    pt1, tt1 = totals(rho1, u1, p1)
    pt2, tt2 = totals(rho2, u2, p2)
    m = 0.5 * rho1 * u1[0] * (a_in + a_out)
    assert abs(t[6] - m) <= 1e-13 * m
    PR, TR = pt2 / pt1, tt2 / tt1
    assert abs(t[7] - PR) <= 1e-13 * PR
    assert abs(t[8] - TR) <= 1e-13 * TR
    dT, dTs = TR - 1.0, PR ** ((G - 1.0) / G) - 1.0
    eta = dTs / dT if dT >= 0.0 else dT / dTs
    assert abs(t[9] - eta) <= 1e-9 * abs(eta)
