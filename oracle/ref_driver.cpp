// Test-only driver around the reference's own path translation units.
//
// TEST INFRASTRUCTURE. This file is compiled together with the unchanged
// reference sources (src/config.cpp, src/input.cpp, src/mesh.cpp,
// src/math_utils.cpp, src/fluid_ideal.cpp, src/advection_{base,jst,roe}.cpp,
// src/boundary_{base,inlet,outlet,outlet_radial_equilibrium,inviscid_wall,
// fake,outlet_throttle,outlet_mass_flow}.cpp, src/greitzer_model.cpp,
// src/source_bfm_{base,hall,lift_drag,thollet}.cpp, src/input_table.cpp,
// src/fluid_real.cpp, src/output.cpp, src/turbulence_model_base.cpp)
// into oracle/_ref/libtbfm_ref.so by oracle/Makefile.  Only tests/ and
// tests/golden/gen_golden.py load it, to pin the CPU restatement in
// oracle/tbfm_oracle.c and the product's host geometry.  src/solver.cpp is not
// part of the build: it needs <Eigen/Sparse>, which this image lacks.
//
// Every entry point below calls a PUBLIC member of a reference class:
//   Mesh getters                        include/mesh.hpp:20-150
//   AdvectionBase::computeFlux          include/advection_base.hpp:21-25
//   BoundaryBase::computeBoundaryFlux   include/boundary_base.hpp:26-32
//   SourceBFMBase::computeTotalSource   include/source_bfm_base.hpp:20-31
//   GreitzerModel                       include/greitzer_model.hpp:7-28
//   FluidIdeal / FluidReal              include/fluid_ideal.hpp, include/fluid_real.hpp
//   Output::writeSolution               include/output.hpp:30-34
#include "types.hpp"
#include "config.hpp"
#include "mesh.hpp"
#include "math_utils.hpp"
#include "fluid_ideal.hpp"
#include "fluid_real.hpp"
#include "advection_jst.hpp"
#include "advection_roe.hpp"
#include "boundary_inlet.hpp"
#include "boundary_outlet.hpp"
#include "boundary_outlet_radial_equilibrium.hpp"
#include "boundary_inviscid_wall.hpp"
#include "boundary_outlet_throttle.hpp"
#include "boundary_outlet_mass_flow.hpp"
#include "greitzer_model.hpp"
#include "source_bfm_hall.hpp"
#include "source_bfm_lift_drag.hpp"
#include "source_bfm_thollet.hpp"
#include "output.hpp"
#include "turbulence_model_none.hpp"

#include <filesystem>

#include <memory>
#include <cstdio>
#include <vector>

namespace {
struct RefCase {
    std::unique_ptr<Config> config;
    std::unique_ptr<Mesh> mesh;
    std::unique_ptr<FluidBase> fluid;
    std::unique_ptr<AdvectionBase> advection;
    std::unique_ptr<SourceBFMBase> bfm;
    std::vector<FloatType> radialProfile;
    // side outputs the reference keeps as Solver members (solver.hpp:370-371), persisting
    // across ref_bfm_sweep calls as they persist across stages
    std::unique_ptr<Matrix3D<Vector3D>> sweepFinv, sweepFvisc;
    std::unique_ptr<Matrix3D<FloatType>> sweepDev, sweepDt;
};
RefCase g_case;

StateVector sv(const double* p) { return StateVector({p[0], p[1], p[2], p[3], p[4]}); }
void put(const StateVector& s, double* out) { for (int v = 0; v < 5; ++v) out[v] = s[v]; }

void dumpVec(const Matrix3D<Vector3D>& m, double* out) {
    const size_t n = m.sizeI() * m.sizeJ() * m.sizeK();
    for (size_t idx = 0; idx < n; ++idx) {
        const Vector3D& v = m[idx];
        out[3 * idx + 0] = v.x();
        out[3 * idx + 1] = v.y();
        out[3 * idx + 2] = v.z();
    }
}
}  // namespace

extern "C" {

// Loads input.ini (absolute paths inside) and builds Config, Mesh, FluidIdeal,
// the configured advection scheme and BFM model.  Returns 0 on success.
int ref_case_load(const char* ini_path) {
    try {
        std::fflush(stdout);
        g_case = RefCase();
        g_case.config = std::make_unique<Config>(std::string(ini_path));
        g_case.mesh = std::make_unique<Mesh>(*g_case.config);
        if (g_case.config->getFluidModel() == FluidModel::REAL)
            g_case.fluid = std::make_unique<FluidReal>(g_case.config->getFluidTableFile());
        else
            g_case.fluid = std::make_unique<FluidIdeal>(g_case.config->getFluidGamma(),
                                                        g_case.config->getFluidGasConstant());
        if (g_case.config->getAdvectionScheme() == AdvectionScheme::ROE)
            g_case.advection = std::make_unique<AdvectionRoe>(*g_case.config, *g_case.fluid);
        else
            g_case.advection = std::make_unique<AdvectionJst>(*g_case.config, *g_case.fluid);
        BodyForceModel m = g_case.config->getBFMModel();
        if (m == BodyForceModel::HALL)
            g_case.bfm = std::make_unique<SourceBFMHall>(*g_case.config, *g_case.fluid, *g_case.mesh);
        else if (m == BodyForceModel::LIFT_DRAG)
            g_case.bfm = std::make_unique<SourceBFMLiftDrag>(*g_case.config, *g_case.fluid, *g_case.mesh);
        else if (m == BodyForceModel::HALL_THOLLET)
            g_case.bfm = std::make_unique<SourceBFMThollet>(*g_case.config, *g_case.fluid, *g_case.mesh);
        else
            g_case.bfm = std::make_unique<SourceBFMBase>(*g_case.config, *g_case.fluid, *g_case.mesh);
        std::fflush(stdout);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "ref_case_load: %s\n", e.what());
        return -1;
    }
    return 0;
}

void ref_dims(long* ni, long* nj, long* nk) {
    *ni = (long)g_case.mesh->getNumberPointsI();
    *nj = (long)g_case.mesh->getNumberPointsJ();
    *nk = (long)g_case.mesh->getNumberPointsK();
}

// which: 0..2 surfaces I/J/K, 3..5 face centres I/J/K, 6 volumes, 7 dual nodes,
// 8 vertices, 9 blockage Green-Gauss gradient (BFM cases only).
// Arrays are in the reference's k-fastest order, Vector3D as xyz triples.
int ref_get_geometry(int which, double* out) {
    const Mesh& m = *g_case.mesh;
    switch (which) {
        case 0: dumpVec(m.getSurfacesI(), out); return 0;
        case 1: dumpVec(m.getSurfacesJ(), out); return 0;
        case 2: dumpVec(m.getSurfacesK(), out); return 0;
        case 3: dumpVec(m.getMidPointsI(), out); return 0;
        case 4: dumpVec(m.getMidPointsJ(), out); return 0;
        case 5: dumpVec(m.getMidPointsK(), out); return 0;
        case 6: {
            const auto& v = m.getVolumes();
            for (size_t i = 0; i < v.totalSize(); ++i) out[i] = v[i];
            return 0;
        }
        case 7: dumpVec(m.getDualNodes(), out); return 0;
        case 8: dumpVec(m.getVertices(), out); return 0;
        case 9: {
            const size_t ni = m.getNumberPointsI(), nj = m.getNumberPointsJ(), nk = m.getNumberPointsK();
            size_t idx = 0;
            for (size_t i = 0; i < ni; ++i)
                for (size_t j = 0; j < nj; ++j)
                    for (size_t k = 0; k < nk; ++k, ++idx) {
                        Vector3D g = m.getInputFieldsGradient(InputField::BLOCKAGE, i, j, k);
                        out[3 * idx + 0] = g.x();
                        out[3 * idx + 1] = g.y();
                        out[3 * idx + 2] = g.z();
                    }
            return 0;
        }
        default: return -1;
    }
}

// computeGradientsGreenGauss4 (math_utils.cpp:237-392) of four node fields (N each,
// k-fastest) over the loaded mesh; out: [field][x, y, z][N].
void ref_gradients4(const double* f, double* out) {
    const Mesh& m = *g_case.mesh;
    const size_t ni = m.getNumberPointsI(), nj = m.getNumberPointsJ(), nk = m.getNumberPointsK(), N = ni * nj * nk;
    Matrix3D<FloatType> F[4];
    Matrix3D<Vector3D> G[4];
    for (int q = 0; q < 4; ++q) {
        F[q].resize(ni, nj, nk);
        G[q].resize(ni, nj, nk);
        for (size_t n = 0; n < N; ++n) F[q][n] = f[q * N + n];
    }
    computeGradientsGreenGauss4(m.getSurfacesI(), m.getSurfacesJ(), m.getSurfacesK(), m.getMidPointsI(),
                                m.getMidPointsJ(), m.getMidPointsK(), m.getVertices(), m.getVolumes(), F[0], F[1],
                                F[2], F[3], G[0], G[1], G[2], G[3]);
    for (int q = 0; q < 4; ++q)
        for (size_t n = 0; n < N; ++n) {
            out[(3 * q) * N + n] = G[q][n].x();
            out[(3 * q + 1) * N + n] = G[q][n].y();
            out[(3 * q + 2) * N + n] = G[q][n].z();
        }
}

// Molecular viscosity of the loaded case's fluid after setTransportProperties
// (fluid_ideal.cpp:209-231), as the solver sets it up when VISCOSITY_ACTIVE (solver.cpp:72-74).
double ref_mu(double T) {
    g_case.fluid->setTransportProperties(*g_case.config);
    return g_case.fluid->computeMolecularDynamicViscosity(T);
}

// computeViscousStressTensor (math_utils.cpp:593-608): out xx, yy, zz, xy, xz, yz.
void ref_stress_tensor(double mu, double lam, const double* gu, const double* gv, const double* gw, double* out) {
    ViscousStressTensor t = computeViscousStressTensor(mu, lam, Vector3D(gu[0], gu[1], gu[2]),
                                                       Vector3D(gv[0], gv[1], gv[2]), Vector3D(gw[0], gw[1], gw[2]));
    out[0] = t.xx; out[1] = t.yy; out[2] = t.zz; out[3] = t.xy; out[4] = t.xz; out[5] = t.yz;
}

// Face flux of the configured scheme (JST or Roe[+MUSCL]), per unit area.
void ref_face_flux(const double* Ull, const double* Ul, const double* Ur, const double* Urr,
                   const double* S, double* out) {
    StateVector f = g_case.advection->computeFlux(sv(Ull), sv(Ul), sv(Ur), sv(Urr),
                                                  Vector3D(S[0], S[1], S[2]));
    put(f, out);
}

// Boundary flux.  kind: 0 inlet, 1 outlet, 2 radial equilibrium (profile given),
// 3 inviscid wall, 4 throttle outlet (profile given), 5 mass-flow outlet (profile given).  values: the patch values of input.ini.  face = {i,j,k} of the
// boundary face as the reference passes it.
void ref_boundary_flux(int kind, const double* values, int nvalues, const double* Uint,
                       const double* S, const long* face, long iter, const double* profile,
                       int nprofile, double* out) {
    std::vector<FloatType> vals(values, values + nvalues);
    std::array<size_t, 3> idx{(size_t)face[0], (size_t)face[1], (size_t)face[2]};
    FlowSolution dummy;
    Vector3D surf(S[0], S[1], S[2]);
    Vector3D mid(0, 0, 0);
    StateVector f;
    if (kind == 0) {
        BoundaryInlet bc(*g_case.config, *g_case.mesh, *g_case.fluid, vals);
        f = bc.computeBoundaryFlux(sv(Uint), surf, mid, idx, dummy, (size_t)iter);
    } else if (kind == 1) {
        BoundaryOutlet bc(*g_case.config, *g_case.mesh, *g_case.fluid, vals);
        f = bc.computeBoundaryFlux(sv(Uint), surf, mid, idx, dummy, (size_t)iter);
    } else if (kind == 2) {
        g_case.radialProfile.assign(profile, profile + nprofile);
        BoundaryOutletRadialEquilibrium bc(*g_case.config, *g_case.mesh, *g_case.fluid,
                                           g_case.radialProfile);
        f = bc.computeBoundaryFlux(sv(Uint), surf, mid, idx, dummy, (size_t)iter);
    } else if (kind == 4) {
        g_case.radialProfile.assign(profile, profile + nprofile);
        BoundaryOutletThrottle bc(*g_case.config, *g_case.mesh, *g_case.fluid, g_case.radialProfile);
        f = bc.computeBoundaryFlux(sv(Uint), surf, mid, idx, dummy, (size_t)iter);
    } else if (kind == 5) {
        g_case.radialProfile.assign(profile, profile + nprofile);
        BoundaryOutletMassFlow bc(*g_case.config, *g_case.mesh, *g_case.fluid, g_case.radialProfile);
        f = bc.computeBoundaryFlux(sv(Uint), surf, mid, idx, dummy, (size_t)iter);
    } else {
        BoundaryInviscidWall bc(*g_case.config, *g_case.mesh, *g_case.fluid);
        f = bc.computeBoundaryFlux(sv(Uint), surf, mid, idx, dummy, (size_t)iter);
    }
    put(f, out);
}

// GreitzerModel (src/greitzer_model.cpp:3-56) of the loaded case's configuration and
// fluid with throttle coefficient kt: initializeState(p0, m_in0, m_out0, T0), then one
// computePlenumPressure per compressor mass flow; pout / mout receive the plenum
// pressure and the plenum outlet mass flow after each step.
void ref_greitzer_steps(double kt, double p0, double min0, double mout0, double T0, const double* mflow, int n,
                        double* pout, double* mout) {
    GreitzerModel gm(*g_case.config, *g_case.fluid, kt);
    gm.initializeState(p0, min0, mout0, T0);
    for (int q = 0; q < n; ++q) {
        pout[q] = gm.computePlenumPressure(mflow[q]);
        mout[q] = gm.getPlenumOutletMassflow((size_t)q + 1);
    }
}

// The body-force part of Solver::computeSourceResiduals (src/solver.cpp:2814-2896) over
// every node in the reference's loop order (i, then j, then k): at nodes whose blockage
// gradient exceeds 1e-10 in magnitude, R -= computeTotalSource(primitives of U), with
// one model object and the force, deviation and time-step arrays persisting across
// calls (Solver members in the reference), as over the stages of a run.  U and R are
// 5*N in the reference layout; finv / fvisc (3*N) and dev (N) receive the side outputs
// after the sweep.
void ref_bfm_sweep(const double* U, double time, double* R, double* finv, double* fvisc, double* dev) {
    const Mesh& m = *g_case.mesh;
    const size_t ni = m.getNumberPointsI(), nj = m.getNumberPointsJ(), nk = m.getNumberPointsK();
    const size_t N = ni * nj * nk;
    if (!g_case.sweepFinv) {
        g_case.sweepFinv = std::make_unique<Matrix3D<Vector3D>>(ni, nj, nk);
        g_case.sweepFvisc = std::make_unique<Matrix3D<Vector3D>>(ni, nj, nk);
        g_case.sweepDev = std::make_unique<Matrix3D<FloatType>>(ni, nj, nk);
        g_case.sweepDt = std::make_unique<Matrix3D<FloatType>>(ni, nj, nk);
    }
    FlowSolution sol(ni, nj, nk);
    for (size_t i = 0; i < ni; ++i)
        for (size_t j = 0; j < nj; ++j)
            for (size_t k = 0; k < nk; ++k) {
                const size_t n = (i * nj + j) * nk + k;
                sol.set(i, j, k, StateVector({U[n], U[N + n], U[2 * N + n], U[3 * N + n], U[4 * N + n]}));
            }
    FloatType t = time, t2 = time;
    for (size_t i = 0; i < ni; ++i)
        for (size_t j = 0; j < nj; ++j)
            for (size_t k = 0; k < nk; ++k) {
                StateVector primitive = getPrimitiveVariablesFromConservative(sol.at(i, j, k));
                Vector3D gb = m.getInputFieldsGradient(InputField::BLOCKAGE, i, j, k);
                if (gb.magnitude() > 1E-10) {
                    StateVector s = g_case.bfm->computeTotalSource(i, j, k, primitive, *g_case.sweepFinv,
                                                                   *g_case.sweepFvisc, *g_case.sweepDev, t, sol,
                                                                   *g_case.sweepDt, t2);
                    const size_t n = (i * nj + j) * nk + k;
                    for (int v = 0; v < 5; ++v) R[v * N + n] -= s[v];
                }
            }
    dumpVec(*g_case.sweepFinv, finv);
    dumpVec(*g_case.sweepFvisc, fvisc);
    for (size_t idx = 0; idx < N; ++idx) dev[idx] = (*g_case.sweepDev)[idx];
}

// Body-force source of the configured BFM model at node (i,j,k) for the given
// primitive state; also returns the inviscid/viscous force vectors and the
// deviation angle the reference writes as side outputs.
void ref_bfm_source(long i, long j, long k, const double* primitive, double time,
                    double* out, double* finv, double* fvisc, double* dev) {
    const Mesh& m = *g_case.mesh;
    const size_t ni = m.getNumberPointsI(), nj = m.getNumberPointsJ(), nk = m.getNumberPointsK();
    Matrix3D<Vector3D> fi(ni, nj, nk), fv(ni, nj, nk);
    Matrix3D<FloatType> da(ni, nj, nk), dt(ni, nj, nk);
    FlowSolution U(ni, nj, nk);
    FloatType t = time, t2 = time;
    StateVector s = g_case.bfm->computeTotalSource((size_t)i, (size_t)j, (size_t)k, sv(primitive),
                                                   fi, fv, da, t, U, dt, t2);
    put(s, out);
    Vector3D a = fi(i, j, k), b = fv(i, j, k);
    finv[0] = a.x(); finv[1] = a.y(); finv[2] = a.z();
    fvisc[0] = b.x(); fvisc[1] = b.y(); fvisc[2] = b.z();
    *dev = da(i, j, k);
}

// Fluid helpers used on the path: out = {p, a, ht, T, pt, Tt, s}.
void ref_fluid_state(const double* primitive, double* out) {
    const FluidBase& f = *g_case.fluid;
    Vector3D u(primitive[1], primitive[2], primitive[3]);
    out[0] = f.computePressure_rho_u_et(primitive[0], u, primitive[4]);
    out[1] = f.computeSoundSpeed_rho_u_et(primitive[0], u, primitive[4]);
    out[2] = f.computeTotalEnthalpy_rho_u_et(primitive[0], u, primitive[4]);
    out[3] = f.computeTemperature_rho_u_et(primitive[0], u, primitive[4]);
    out[4] = f.computeTotalPressure_rho_u_et(primitive[0], u, primitive[4]);
    out[5] = f.computeTotalTemperature_rho_u_et(primitive[0], u, primitive[4]);
    out[6] = f.computeEntropy_rho_u_et(primitive[0], u, primitive[4]);
}

// Point functions of the loaded fluid (public FluidBase members), numbered as the
// oracle's tbo_lut_eval: 0 p(rho,e) 1 T(rho,e) 2 a(rho,e) 3 s(rho,e) 4 dp/drho|e
// 5 dp/de|rho 6 e(p,rho) 7 a(p,rho) 8 rho(p,T) 9 s(p,T) 10 rho(p,s) 11 T(p,s) 12 e(p,s)
// 13 h(p,s) 14 e(p,T) 15 a(p,T) 16 p(rho,T).
double ref_fluid_eval(int which, double x, double y) {
    const FluidBase& f = *g_case.fluid;
    switch (which) {
        case 0: return f.computePressure_rho_e(x, y);
        case 1: return f.computeTemperature_rho_e(x, y);
        case 2: return f.computeSoundSpeed_rho_e(x, y);
        case 3: return f.computeEntropy_rho_e(x, y);
        case 4: return f.computeDpDrho_e(x, y);
        case 5: return f.computeDpDe_rho(x, y);
        case 6: return f.computeStaticEnergy_p_rho(x, y);
        case 7: return f.computeSoundSpeed_p_rho(x, y);
        case 8: return f.computeDensity_p_T(x, y);
        case 9: return f.computeEntropy_p_T(x, y);
        case 10: return f.computeDensity_p_s(x, y);
        case 11: return f.computeTemperature_p_s(x, y);
        case 12: return f.computeInternalEnergy_p_s(x, y);
        case 13: return f.computeEnthalpy_p_s(x, y);
        case 14: return f.computeInternalEnergy_p_T(x, y);
        case 15: return f.computeSoundSpeed_p_T(x, y);
        case 16: return f.computePressure_rho_T(x, y);
        default: return 0.0;
    }
}

// Initial free-stream state of computeInitFields (fluid_ideal.cpp / fluid_real.cpp).
void ref_init_state(double mach, double T, double p, const double* dir, double* U) {
    FloatType rho = 0, et = 0;
    Vector3D vel;
    g_case.fluid->computeInitFields(mach, T, p, Vector3D(dir[0], dir[1], dir[2]), rho, vel, et);
    U[0] = rho; U[1] = rho * vel.x(); U[2] = rho * vel.y(); U[3] = rho * vel.z(); U[4] = rho * et;
}

// Trapezoidal radial-equilibrium integration (math_utils.cpp:141-170).
void ref_radial_equilibrium(const double* rho, const double* ut, const double* r, int n,
                            double hub, double* p) {
    std::vector<FloatType> a(rho, rho + n), b(ut, ut + n), c(r, r + n), out(n);
    FloatType h = hub;
    integrateRadialEquilibrium(a, b, c, out, h);
    for (int i = 0; i < n; ++i) p[i] = out[i];
}

// Volume solution file through the reference's own writer: Output::writeSolution
// (src/output.cpp:367-447) in folder `dir` (Volume_CSV/<SOLUTION_NAME>[_<it>].csv) for
// the state U (5 x N, k fastest) and side outputs finv / fvisc (xyz triples) and
// devang.  No boundaries, no turbulence model (TurbulenceModelNone), as on the Euler
// path.  Returns 0 on success.
int ref_write_solution(const double* U, const double* finv, const double* fvisc, const double* devang, long it,
                       const char* dir) {
    try {
        const Mesh& m = *g_case.mesh;
        const size_t ni = m.getNumberPointsI(), nj = m.getNumberPointsJ(), nk = m.getNumberPointsK();
        const size_t N = ni * nj * nk;
        FlowSolution sol(ni, nj, nk);
        Matrix3D<Vector3D> inv(ni, nj, nk), visc(ni, nj, nk);
        Matrix3D<FloatType> dev(ni, nj, nk), wall(ni, nj, nk);
        for (size_t n = 0; n < N; ++n) {
            sol._rho[n] = U[n];
            sol._rhoU[n] = U[N + n];
            sol._rhoV[n] = U[2 * N + n];
            sol._rhoW[n] = U[3 * N + n];
            sol._rhoE[n] = U[4 * N + n];
            inv[n] = Vector3D(finv[3 * n], finv[3 * n + 1], finv[3 * n + 2]);
            visc[n] = Vector3D(fvisc[3 * n], fvisc[3 * n + 1], fvisc[3 * n + 2]);
            dev[n] = devang[n];
        }
        std::map<SolutionName, Matrix3D<Vector3D>> grad;
        std::vector<Boundary> boundaries;
        TurbulenceModelNone turb(*g_case.config, *g_case.fluid, m, boundaries, wall, sol);
        const std::filesystem::path here = std::filesystem::current_path();
        std::filesystem::current_path(dir);
        {
            Output out(*g_case.config, m, sol, grad, *g_case.fluid, boundaries, turb, inv, visc, dev, wall);
            out.writeSolution((size_t)it);
        }
        std::filesystem::current_path(here);
        std::fflush(stdout);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "ref_write_solution: %s\n", e.what());
        return -1;
    }
    return 0;
}

}  // extern "C"
